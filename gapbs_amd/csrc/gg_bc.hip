// Brandes betweenness centrality on gfx950.
//
// Reference: Brandes / PBFS (src/bc.cc:51-150).  Kept, per root:
//   * a level-synchronous top-down BFS claiming depths with a compare-and-
//     swap (bc.cc:65-77); the reached vertices in BFS order with per-level
//     boundaries (the reference's queue and depth_index, bc.cc:56-87);
//   * shortest-path counts sigma in double (bc.cc:47, 78-80);
//   * successors of u are the out-neighbours one level deeper -- the
//     reference records them in an edge bitmap during the BFS (bc.cc:76),
//     here they are re-derived from depths, as its verifier does
//     (bc.cc:207);
//   * for depths from the deepest down to 0 (the root included, whose own
//     dependency is added to its score): delta_u = sum over successors v of
//     (sigma_u / sigma_v) * (1 + delta_v), the double term added into a
//     float delta_u (bc.cc:123-136), then score_u += delta_u;
//   * after all roots, every score divided by the largest (bc.cc:141-148).
// GPU design:
//   * discovery expands the level with gg_expand.h and appends the next
//     level to the same queue through an LDS-staged queue;
//   * sigma is pulled: each newly discovered vertex sums sigma over its
//     in-neighbours at the previous level (no floating-point atomics; the
//     sums are exact integers below 2^53, so the order does not matter);
//   * the sigma pull runs one thread per vertex; vertices with >= 256
//     in-edges are cut into chunks summed by whole blocks and combined with
//     one double atomic per chunk;
//   * delta runs one thread per vertex below 64 edges and one wavefront per
//     vertex above: the wavefront computes 64 successor terms at a time and
//     folds them into the accumulator in successor order (ballot, then one
//     lane after another), so every vertex's delta is rounded exactly as the
//     reference rounds it;
//   * GG_BC_FP32: the reference's float arithmetic throughout (its verifier
//     allows only FLT_EPSILON); GG_BC_FP32_KAHAN: float dependencies and
//     scores, each per-vertex dependency sum and each score sum over roots
//     carrying a float compensation term; GG_BC_FP64: double dependencies
//     and scores.
#include "gg_common.h"
#include "gg_expand.h"

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kQueueCap = 2048;
constexpr int kStripes = 64;
constexpr int64_t kHubEdges = 256;   // sigma pull: vertices with this many in-edges use chunks
constexpr int64_t kWaveEdges = 64;   // delta: vertices with this many edges use a wavefront
constexpr int64_t kBcHubEdges = 16384;  // ... and with this many, a block (BcDeltaHub)
constexpr int kBcHubProducers = 15;
constexpr int kBcHubUnroll = 2;  // 64-edge runs per producer wavefront and round
constexpr int kBcHubGrid = 2;    // BcDeltaHub blocks per CU
constexpr int kBcHubBS = (kBcHubProducers + 1) * kWave;

struct BcCounters {
  int32_t tail;     // queue length (all levels so far)
  int32_t nchunks;  // discovery chunks
  int32_t nred;     // sigma reduction chunks
  int32_t nwave;    // delta: vertices handed to the wavefront kernel
  int32_t nhub;     // delta: vertices handed to the hub kernel
  int32_t hub_next; // BcDeltaHub: next hub list entry to take
  unsigned long long examined[kStripes];
};

__device__ __forceinline__ int32_t Load(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBS) void BcResetSource(int32_t* __restrict__ depth, int64_t n, int32_t src,
                                                     double* __restrict__ sigma, int32_t* __restrict__ queue,
                                                     BcCounters* c) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    depth[i] = i == src ? 0 : -1;
  if (blockIdx.x == 0) {
    if (threadIdx.x < kStripes) c->examined[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      sigma[src] = 1.0;
      queue[0] = src;
      c->tail = 1;
      c->nchunks = 0;
      c->nred = 0;
      c->nwave = 0;
      c->nhub = 0;
      c->hub_next = 0;
    }
  }
}

struct DiscoverVisit {
  const int32_t* dst;
  int32_t* depth;
  int32_t next_depth;
  LdsQueue<kQueueCap>* lq;
  int32_t* queue;
  int32_t* tail;
  __device__ __forceinline__ void operator()(int32_t, int32_t, int64_t e, bool valid) {
    bool won = false;
    int32_t v = 0;
    if (valid) {
      v = dst[e];
      won = Load(&depth[v]) == -1 && atomicCAS(&depth[v], -1, next_depth) == -1;
    }
    LdsQueuePush(*lq, won, v, queue, tail);
  }
};

// Claims the next level's vertices and appends them to the queue.
__global__ __launch_bounds__(kBS) void BcDiscover(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                  int32_t* depth, int32_t* queue, int32_t qb, int32_t qn,
                                                  int32_t next_depth, BcCounters* c, EdgeChunk* chunks,
                                                  int64_t chunk_min) {
  __shared__ ExpandShared sh;
  __shared__ LdsQueue<kQueueCap> lq;
  __shared__ unsigned long long s_red[kBS / kWave];
  LdsQueueInit(lq);
  DiscoverVisit visit{dst, depth, next_depth, &lq, queue, &c->tail};
  unsigned long long ex = 0;
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < qn; tile += int64_t(gridDim.x) * kBS) {
    const int64_t i = tile + threadIdx.x;
    int64_t beg = 0, end = 0;
    if (i < qn) {
      const int32_t u = queue[qb + i];
      beg = off[u];
      end = off[u + 1];
      ex += static_cast<unsigned long long>(end - beg);
    }
    ExpandTile(sh, 0, 0, beg, end, visit, chunks, &c->nchunks, chunk_min);
  }
  LdsQueueFlush(lq, queue, &c->tail);
  ex = BlockSum<kBS>(ex, s_red);
  if (threadIdx.x == 0 && ex) atomicAdd(&c->examined[blockIdx.x & (kStripes - 1)], ex);
}

__global__ __launch_bounds__(kBS) void BcDiscoverChunks(const int32_t* __restrict__ dst, int32_t* depth,
                                                        int32_t* queue, int32_t next_depth, BcCounters* c,
                                                        const EdgeChunk* chunks) {
  __shared__ LdsQueue<kQueueCap> lq;
  LdsQueueInit(lq);
  DiscoverVisit visit{dst, depth, next_depth, &lq, queue, &c->tail};
  ExpandChunks(chunks, c->nchunks, visit);
  LdsQueueFlush(lq, queue, &c->tail);
}

// Appends the chunks of a hub vertex (>= kHubEdges edges) for a reduction.
__device__ __forceinline__ void EmitReduceChunks(int32_t u, int64_t beg, int64_t end, EdgeChunk* chunks,
                                                 int32_t* nchunks) {
  const int nc = static_cast<int>((end - beg + kChunkEdges - 1) / kChunkEdges);
  const int base = atomicAdd(nchunks, nc);
  for (int k = 0; k < nc; k++) {
    const int64_t cb = beg + k * kChunkEdges;
    chunks[base + k] = EdgeChunk{cb, cb + kChunkEdges < end ? cb + kChunkEdges : end, u, k == 0};
  }
}

// sigma[v] for the new level: sum of sigma over in-neighbours one level up.
// Hubs get sigma = 0 here and their chunks summed by BcSigmaChunks.
__global__ __launch_bounds__(kBS) void BcSigma(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                               const int32_t* __restrict__ depth, double* __restrict__ sigma,
                                               const int32_t* __restrict__ queue, int32_t qb, int32_t qn,
                                               int32_t prev_depth, EdgeChunk* chunks, BcCounters* c,
                                               unsigned long long* work) {
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long ent = 0, mat = 0;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < qn; i += int64_t(gridDim.x) * kBS) {
    const int32_t v = queue[qb + i];
    const int64_t b = in_off[v], e = in_off[v + 1];
    if (e - b >= kHubEdges) {
      sigma[v] = 0.0;
      EmitReduceChunks(v, b, e, chunks, &c->nred);
      continue;
    }
    double s = 0.0;
    for (int64_t k = b; k < e; k++) {
      const int32_t u = in_dst[k];
      if (depth[u] == prev_depth) {
        s += sigma[u];
        mat++;
      }
    }
    ent += static_cast<unsigned long long>(e - b);
    sigma[v] = s;
  }
  WorkAdd(work, BlockSum<kBS>(ent, s_red), BlockSum<kBS>(mat, s_red));
}

__global__ __launch_bounds__(kBS) void BcSigmaChunks(const int32_t* __restrict__ in_dst,
                                                     const int32_t* __restrict__ depth, double* sigma,
                                                     int32_t prev_depth, const EdgeChunk* chunks,
                                                     const BcCounters* c, unsigned long long* work) {
  __shared__ double s_red[kBS / kWave];
  __shared__ unsigned long long s_cnt[kBS / kWave];
  const int32_t nc = c->nred;
  unsigned long long ent = 0, mat = 0;
  for (int32_t ci = blockIdx.x; ci < nc; ci += gridDim.x) {
    const EdgeChunk ch = chunks[ci];
    double s = 0.0;
    for (int64_t k = ch.beg + threadIdx.x; k < ch.end; k += kBS) {
      const int32_t u = in_dst[k];
      ent++;
      if (depth[u] == prev_depth) {
        s += sigma[u];
        mat++;
      }
    }
    s = BlockSum<kBS>(s, s_red);
    if (threadIdx.x == 0 && s != 0.0) atomicAdd(&sigma[ch.u], s);
  }
  WorkAdd(work, BlockSum<kBS>(ent, s_cnt), BlockSum<kBS>(mat, s_cnt));
}

// delta for one level, the reference's way: a double term added into a
// float (or double) accumulator, in successor order.
__device__ __forceinline__ double DeltaTerm(double sig_u, double sig_v, float delta_v) {
  return __dmul_rn(__ddiv_rn(sig_u, sig_v), static_cast<double>(__fadd_rn(1.0f, delta_v)));
}
__device__ __forceinline__ double DeltaTerm(double sig_u, double sig_v, double delta_v) {
  return __dmul_rn(__ddiv_rn(sig_u, sig_v), __dadd_rn(1.0, delta_v));
}
// Per-vertex dependency accumulators.  RefAcc rounds like the reference (a
// double term added into a float, or double, total); KahanAcc keeps a float
// sum plus a float compensation, fed with each double term split into a
// float head and a float tail.
template <class T>
struct RefAcc {
  T v;
  __device__ __forceinline__ static RefAcc Zero() { return RefAcc{T(0)}; }
  __device__ __forceinline__ void Add(double t) { v = static_cast<T>(__dadd_rn(static_cast<double>(v), t)); }
  __device__ __forceinline__ T Value() const { return v; }
};

struct KahanAcc {
  float s, c;
  __device__ __forceinline__ static KahanAcc Zero() { return KahanAcc{0.0f, 0.0f}; }
  __device__ __forceinline__ void AddF(float x) {  // Neumaier
    const float t = __fadd_rn(s, x);
    const float e = fabsf(s) >= fabsf(x) ? __fadd_rn(__fsub_rn(s, t), x) : __fadd_rn(__fsub_rn(x, t), s);
    c = __fadd_rn(c, e);
    s = t;
  }
  __device__ __forceinline__ void Add(double t) {
    const float hi = static_cast<float>(t);
    const float lo = static_cast<float>(__dsub_rn(t, static_cast<double>(hi)));
    AddF(hi);
    c = __fadd_rn(c, lo);
  }
  __device__ __forceinline__ float Value() const { return __fadd_rn(s, c); }
};

template <class T, bool kKahan>
struct AccFor {
  using Type = RefAcc<T>;
};
template <>
struct AccFor<float, true> {
  using Type = KahanAcc;
};

// score += delta: plain, or with a Kahan compensation term (float only).
template <bool kKahan>
__device__ __forceinline__ void AddScore(float* score, float* comp, int32_t u, float d) {
  if (kKahan) {
    const float y = __fsub_rn(d, comp[u]);
    const float s = score[u];
    const float t = __fadd_rn(s, y);
    comp[u] = __fsub_rn(__fsub_rn(t, s), y);
    score[u] = t;
  } else {
    score[u] = __fadd_rn(score[u], d);
  }
}
template <bool kKahan>
__device__ __forceinline__ void AddScore(double* score, double*, int32_t u, double d) {
  score[u] = __dadd_rn(score[u], d);
}

// Thread per vertex below kWaveEdges edges; larger vertices are listed for
// BcDeltaWave, and those with kBcHubEdges or more for BcDeltaHub.
template <class T, bool kKahan>
__global__ __launch_bounds__(kBS) void BcDelta(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                               const int32_t* __restrict__ depth,
                                               const double* __restrict__ sigma, T* __restrict__ delta,
                                               T* __restrict__ score, T* __restrict__ comp,
                                               const int32_t* __restrict__ queue, int32_t qb, int32_t qn,
                                               int32_t level, int32_t* __restrict__ wave_list,
                                               int32_t* __restrict__ hub_list, BcCounters* c,
                                               unsigned long long* work) {
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long ent = 0, mat = 0;
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < qn; base += int64_t(gridDim.x) * kBS) {
    const int64_t i = base + threadIdx.x;
    bool big = false, hub = false;
    int32_t u = 0;
    if (i < qn) {
      u = queue[qb + i];
      const int64_t b = off[u], e = off[u + 1];
      big = e - b >= kWaveEdges;
      hub = e - b >= kBcHubEdges;
      if (!big) {
        const double su = sigma[u];
        typename AccFor<T, kKahan>::Type acc = AccFor<T, kKahan>::Type::Zero();
        for (int64_t k = b; k < e; k++) {
          const int32_t v = dst[k];
          if (depth[v] == level + 1) {
            acc.Add(DeltaTerm(su, sigma[v], delta[v]));
            mat++;
          }
        }
        ent += static_cast<unsigned long long>(e - b);
        const T d = acc.Value();
        delta[u] = d;
        AddScore<kKahan>(score, comp, u, d);
      }
    }
    WaveAppend(big && !hub, u, wave_list, &c->nwave);
    WaveAppend(hub, u, hub_list, &c->nhub);
  }
  WorkAdd(work, BlockSum<kBS>(ent, s_red), BlockSum<kBS>(mat, s_red));
}

// One wavefront per listed vertex: 64 successor terms computed at once,
// folded into the accumulator in edge order.
template <class T, bool kKahan>
__global__ __launch_bounds__(kBS) void BcDeltaWave(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                   const int32_t* __restrict__ depth,
                                                   const double* __restrict__ sigma, T* __restrict__ delta,
                                                   T* __restrict__ score, T* __restrict__ comp,
                                                   const int32_t* __restrict__ wave_list, int32_t level,
                                                   const BcCounters* c, unsigned long long* work) {
  __shared__ unsigned long long s_red[kBS / kWave];
  const int lane = LaneId();
  unsigned long long ent = 0, mat = 0;  // lane 0's
  const int32_t nw = c->nwave;
  const int64_t wave = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kBS / kWave);
  for (int64_t i = wave; i < nw; i += nwaves) {
    const int32_t u = wave_list[i];
    const int64_t b = off[u], e = off[u + 1];
    const double su = sigma[u];
    typename AccFor<T, kKahan>::Type acc = AccFor<T, kKahan>::Type::Zero();
    for (int64_t base = b; base < e; base += kWave) {
      const int64_t k = base + lane;
      double term = 0.0;
      bool has = false;
      if (k < e) {
        const int32_t v = dst[k];
        if (depth[v] == level + 1) {
          term = DeltaTerm(su, sigma[v], delta[v]);
          has = true;
        }
      }
      uint64_t m = __ballot(has);
      if (lane == 0) mat += static_cast<unsigned long long>(__popcll(m));
      while (m) {
        const int j = __ffsll(static_cast<unsigned long long>(m)) - 1;
        acc.Add(__shfl(term, j));
        m &= m - 1;
      }
    }
    if (lane == 0) {
      ent += static_cast<unsigned long long>(e - b);
      const T d = acc.Value();
      delta[u] = d;
      AddScore<kKahan>(score, comp, u, d);
    }
  }
  WorkAdd(work, BlockSum<kBS>(ent, s_red), BlockSum<kBS>(mat, s_red));
}

// The hub list's vertices (at least kBcHubEdges successors), one block each:
// per round, producer wavefront p computes the terms of 64 kBcHubUnroll
// consecutive edges (from b + 64 kBcHubUnroll (round kBcHubProducers + p))
// and packs the matched ones, in edge order, into one half of an LDS double
// buffer; the last wavefront folds the previous round's terms in edge
// order.  Same terms, same order, same arithmetic as BcDeltaWave; only the
// loads and divisions run in parallel.  The fold (widen, add in double,
// round to float per term, as the reference's float `+=` of a double term)
// is a chain of three dependent double-precision operations per matched
// edge, so the largest hubs set the kernel's time.  Blocks take hubs from
// a counter, so a block held by a large hub takes no others meanwhile (the
// list follows the level's queue, which on a degree-ordered layout runs
// roughly from the largest hub down).
template <class T, bool kKahan>
__global__ __launch_bounds__(kBcHubBS) void BcDeltaHub(const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ dst,
                                                       const int32_t* __restrict__ depth,
                                                       const double* __restrict__ sigma, T* __restrict__ delta,
                                                       T* __restrict__ score, T* __restrict__ comp,
                                                       const int32_t* __restrict__ hub_list, int32_t level,
                                                       BcCounters* c, unsigned long long* work) {
  __shared__ double terms[2][kBcHubProducers][kWave * kBcHubUnroll];
  __shared__ int counts[2][kBcHubProducers];
  __shared__ int32_t s_take;
  __shared__ unsigned long long s_red[kBcHubBS / kWave];
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  unsigned long long ent = 0, mat = 0;  // lane 0 of each producer's
  const int32_t nh = c->nhub;
  constexpr int64_t kRound = int64_t(kBcHubProducers) * kWave * kBcHubUnroll;
  for (;;) {
    if (threadIdx.x == 0) s_take = atomicAdd(&c->hub_next, 1);
    __syncthreads();
    const int32_t i = s_take;
    __syncthreads();  // read by all before the next take overwrites it
    if (i >= nh) break;
    const int32_t u = hub_list[i];
    const int64_t b = off[u], e = off[u + 1];
    const double su = sigma[u];
    typename AccFor<T, kKahan>::Type acc = AccFor<T, kKahan>::Type::Zero();
    const int64_t rounds = (e - b + kRound - 1) / kRound;
    for (int64_t k = 0; k <= rounds; k++) {
      if (wid < kBcHubProducers) {
        if (k < rounds) {
          // kBcHubUnroll consecutive runs of 64 edges, their loads issued together.
          const int64_t k0 = b + k * kRound + int64_t(wid) * (kWave * kBcHubUnroll) + lane;
          int32_t v[kBcHubUnroll];
          bool has[kBcHubUnroll];
#pragma unroll
          for (int r = 0; r < kBcHubUnroll; r++) v[r] = k0 + r * kWave < e ? dst[k0 + r * kWave] : -1;
#pragma unroll
          for (int r = 0; r < kBcHubUnroll; r++) has[r] = v[r] >= 0 && depth[v[r]] == level + 1;
          double term[kBcHubUnroll];
#pragma unroll
          for (int r = 0; r < kBcHubUnroll; r++) term[r] = has[r] ? DeltaTerm(su, sigma[v[r]], delta[v[r]]) : 0.0;
          int pos = 0;
#pragma unroll
          for (int r = 0; r < kBcHubUnroll; r++) {
            const uint64_t m = __ballot(has[r]);
            if (has[r]) terms[k & 1][wid][pos + __popcll(LanesBelow(m))] = term[r];
            pos += __popcll(m);
            const uint64_t in = __ballot(k0 + r * kWave < e);
            if (lane == 0) ent += static_cast<unsigned long long>(__popcll(in));
          }
          if (lane == 0) {
            counts[k & 1][wid] = pos;
            mat += static_cast<unsigned long long>(pos);
          }
        }
      } else if (k >= 1) {
        const int h = (k - 1) & 1;
        for (int p = 0; p < kBcHubProducers; p++) {
          const int cnt = counts[h][p];
          int j = 0;
          for (; j + 8 <= cnt; j += 8) {  // eight LDS reads in flight, then the ordered fold
            double t[8];
#pragma unroll
            for (int q = 0; q < 8; q++) t[q] = terms[h][p][j + q];
#pragma unroll
            for (int q = 0; q < 8; q++) acc.Add(t[q]);
          }
          for (; j < cnt; j++) acc.Add(terms[h][p][j]);
        }
      }
      __syncthreads();
    }
    if (wid == kBcHubProducers && lane == 0) {
      const T d = acc.Value();
      delta[u] = d;
      AddScore<kKahan>(score, comp, u, d);
    }
  }
  WorkAdd(work, BlockSum<kBcHubBS>(ent, s_red), BlockSum<kBcHubBS>(mat, s_red));
}


// ---- discovery by DOBFS ------------------------------------------------------
//
// Depths are unique, so the direction-optimizing BFS (RunDobfs) gives the
// levels of the reference's top-down PBFS at a fraction of the edges.  The
// vertices are then bucketed by depth into the queue layout the sigma and
// dependency passes read (level d at queue[bounds[d], bounds[d + 1]); the
// order inside a level does not affect any result).
constexpr int kBucketLevels = 4096;  // deeper traversals use the top-down discovery
constexpr int64_t kBucketChunk = 65536;

__global__ __launch_bounds__(kBS) void BcLevelCount(const int32_t* __restrict__ depth, int64_t n, int nlev,
                                                    int32_t* counts) {
  __shared__ int32_t h[kBucketLevels];
  for (int i = threadIdx.x; i < nlev; i += kBS) h[i] = 0;
  __syncthreads();
  const int64_t lo = int64_t(blockIdx.x) * kBucketChunk;
  const int64_t hi = lo + kBucketChunk < n ? lo + kBucketChunk : n;
  for (int64_t v = lo + threadIdx.x; v < hi; v += kBS) {
    const int32_t d = depth[v];
    if (d >= 0 && d < nlev) atomicAdd(&h[d], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nlev; i += kBS)
    if (h[i]) atomicAdd(&counts[i], h[i]);
}

// cursor[d] starts at bounds[d]; each block reserves its ranges with one
// atomic per level it holds.
__global__ __launch_bounds__(kBS) void BcLevelScatter(const int32_t* __restrict__ depth, int64_t n, int nlev,
                                                      int32_t* cursor, int32_t* __restrict__ queue) {
  __shared__ int32_t h[kBucketLevels];
  for (int i = threadIdx.x; i < nlev; i += kBS) h[i] = 0;
  __syncthreads();
  const int64_t lo = int64_t(blockIdx.x) * kBucketChunk;
  const int64_t hi = lo + kBucketChunk < n ? lo + kBucketChunk : n;
  for (int64_t v = lo + threadIdx.x; v < hi; v += kBS) {
    const int32_t d = depth[v];
    if (d >= 0 && d < nlev) atomicAdd(&h[d], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nlev; i += kBS)
    if (h[i]) h[i] = atomicAdd(&cursor[i], h[i]);
  __syncthreads();
  for (int64_t v = lo + threadIdx.x; v < hi; v += kBS) {
    const int32_t d = depth[v];
    if (d >= 0 && d < nlev) queue[atomicAdd(&h[d], 1)] = static_cast<int32_t>(v);
  }
}

__global__ __launch_bounds__(kBS) void BcClearReduce(BcCounters* c) {
  if (threadIdx.x == 0) {
    c->nred = 0;
    c->nwave = 0;
    c->nhub = 0;
    c->hub_next = 0;
  }
}

enum { kMailTail, kMailExamined, kMailValues };

__global__ __launch_bounds__(kBS) void BcPublish(BcCounters* c, unsigned long long* mb, unsigned long long seq) {
  if (threadIdx.x == 0) {
    unsigned long long vals[kMailValues];
    unsigned long long ex = 0;
    for (int i = 0; i < kStripes; i++) {
      ex += c->examined[i];
      c->examined[i] = 0;
    }
    vals[kMailTail] = static_cast<unsigned long long>(c->tail);
    vals[kMailExamined] = ex;
    c->nchunks = 0;
    c->nred = 0;
    MailPost(mb, vals, kMailValues, seq);
  }
}

template <class T>
__global__ __launch_bounds__(kBS) void BcZero(T* __restrict__ score, T* __restrict__ comp, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    score[i] = 0;
    if (comp) comp[i] = 0;
  }
}

// Largest score (scores are >= 0, so their bit patterns order like the
// values), then every score divided by it (bc.cc:141-148).
template <class T>
__device__ __forceinline__ T Final(const T* score, const T* comp, int64_t i);
template <>
__device__ __forceinline__ float Final(const float* score, const float* comp, int64_t i) {
  return __fadd_rn(score[i], -comp[i]);
}
template <>
__device__ __forceinline__ double Final(const double* score, const double*, int64_t i) {
  return score[i];
}

template <class T>
__global__ __launch_bounds__(kBS) void BcMax(const T* __restrict__ score, const T* __restrict__ comp, int64_t n,
                                             unsigned long long* max_bits) {
  __shared__ double s_max[kBS / kWave];
  double m = 0.0;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    const double v = static_cast<double>(Final(score, comp, i));
    m = v > m ? v : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(m, o);
    m = t > m ? t : m;
  }
  if (LaneId() == 0) s_max[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBS / kWave; w++) m = s_max[w] > m ? s_max[w] : m;
    atomicMax(max_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

template <class T>
__global__ __launch_bounds__(kBS) void BcNormalize(const T* __restrict__ score, const T* __restrict__ comp,
                                                   int64_t n, const unsigned long long* max_bits,
                                                   float* __restrict__ out) {
  const double mx = __longlong_as_double(static_cast<long long>(*max_bits));
  const T biggest = static_cast<T>(mx);
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    out[i] = static_cast<float>(Final(score, comp, i) / biggest);
}

template <class T, bool kKahan>
void RunBrandes(gg_graph* g, const int32_t* roots, int k, float* scores_out, gg_stats* stats,
                const WallTimer& wt) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  const char* tag = sizeof(T) == 4 ? "32" : "64";
  int32_t* depth = g->Scratch<int32_t>("bc.depth", n);
  double* sigma = g->Scratch<double>("bc.sigma", n);
  int32_t* queue = g->Scratch<int32_t>("bc.queue", n + 64);
  T* delta = g->Scratch<T>((std::string("bc.delta") + tag).c_str(), n);
  T* score = g->Scratch<T>((std::string("bc.score") + tag).c_str(), n);
  T* comp = sizeof(T) == 4 ? g->Scratch<T>("bc.comp32", n) : nullptr;  // zero unless kKahan
  int32_t* wave_list = g->Scratch<int32_t>("bc.wave_list", n + 64);
  // A vertex with kBcHubEdges successors owns that many list entries.
  int32_t* hub_list = g->Scratch<int32_t>("bc.hub_list", g->out_entries / kBcHubEdges + 64);
  float* out32 = g->Scratch<float>("bc.out32", n);
  unsigned long long* max_bits = g->Scratch<unsigned long long>("bc.max", 1);
  BcCounters* dc = g->Scratch<BcCounters>("bc.ctr", 1);
  // Work counters (entries read, matched in-neighbours / successors) per
  // kernel kind, for the profiled byte credits.
  enum { kWSigma, kWSigmaChunks, kWDelta, kWDeltaWave, kWDeltaHub, kWKinds };
  unsigned long long* work = g->Scratch<unsigned long long>("bc.work", kWKinds * kWorkCounters);
  int rec[kWKinds] = {-1, -1, -1, -1, -1};
  auto note = [&](int kind, int idx) {
    if (rec[kind] < 0) rec[kind] = idx;
  };
  auto wk = [&](int kind) { return work + kind * kWorkCounters; };
  int32_t* lvl_counts = g->Scratch<int32_t>("bc.level_counts", kBucketLevels);
  int32_t* lvl_cursor = g->Scratch<int32_t>("bc.level_cursor", kBucketLevels);
  std::vector<int32_t> h_counts(kBucketLevels);
  const int64_t entries = g->out_entries > g->in_entries ? g->out_entries : g->in_entries;
  EdgeChunk* chunks = g->Scratch<EdgeChunk>("bc.chunks", MaxChunks(n, entries));
  EdgeChunk* rchunks = g->Scratch<EdgeChunk>("bc.rchunks", MaxChunks(n, entries));
  const bool hubs = g->max_degree >= kHubEdges;
  const dim3 blk(kBS);
  const dim3 grid_n(GridFor(n, kBS, g->cu_count * 32));
  const dim3 grid_c(g->cu_count * 4);
  const double N = static_cast<double>(n);
  int64_t launches = 0, levels_total = 0, reached_total = 0, examined_total = 0;
  unsigned long long mail[kMailValues];

  // Dependencies are summed in each vertex's uploaded list order (the
  // reference's successor order), also on a degree-ordered handle.
  const int32_t* fold_dst = UploadOrderLists(g);
  BeginTiming(g);
  Run(g, "bc_zero", N * sizeof(T) * 2, BcZero<T>, grid_n, blk, score, comp, n);
  Run(g, "bc_zero", 0, WorkClear, dim3(1), dim3(256), work, kWKinds * kWorkCounters);
  launches += 2;
  std::vector<int32_t> bounds;
  for (int r = 0; r < k; r++) {
    Run(g, "bc_reset_source", 4 * N, BcResetSource, grid_n, blk, depth, n, roots[r], sigma, queue, dc);
    launches++;
    bounds.assign(1, 0);
    bounds.push_back(1);
    int32_t level = 0;
    // Forward, by DOBFS and bucketing when the levels fit the buckets.
    const BfsRun br = RunDobfs(g, roots[r], 15, 18, depth);
    launches += br.launches;
    examined_total += br.examined;
    const int nlev = static_cast<int>(br.td + br.bu + 1);
    bool bucketed = false;
    if (nlev <= kBucketLevels) {
      GG_HIP(hipMemsetAsync(lvl_counts, 0, sizeof(int32_t) * nlev, st));
      const dim3 grid_b(GridFor(n, static_cast<int>(kBucketChunk), 1u << 20));
      Run(g, "bc_level_count", 4 * N, BcLevelCount, grid_b, blk, static_cast<const int32_t*>(depth), n, nlev,
          lvl_counts);
      GG_HIP(hipMemcpyAsync(h_counts.data(), lvl_counts, sizeof(int32_t) * nlev, hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      int last = 0;
      for (int d = 0; d < nlev; d++)
        if (h_counts[d]) last = d;
      bounds.assign(1, 0);
      for (int d = 0; d <= last; d++) bounds.push_back(bounds.back() + h_counts[d]);
      GG_HIP(hipMemcpyAsync(lvl_cursor, bounds.data(), sizeof(int32_t) * (last + 1), hipMemcpyHostToDevice, st));
      Run(g, "bc_level_scatter", 8 * N, BcLevelScatter, grid_b, blk, static_cast<const int32_t*>(depth), n, last + 1,
          lvl_cursor, queue);
      launches += 2;
      for (int32_t d = 1; d <= last; d++) {
        const int32_t nb = bounds[d], nn = bounds[d + 1] - bounds[d];
        note(kWSigma, Run(g, "bc_sigma", 20.0 * nn, BcSigma, dim3(GridFor(nn, kBS, g->cu_count * 16)), blk,
                          static_cast<const int64_t*>(g->in_off), static_cast<const int32_t*>(g->in_dst),
                          static_cast<const int32_t*>(depth), sigma, static_cast<const int32_t*>(queue), nb, nn,
                          d - 1, rchunks, dc, wk(kWSigma)));
        launches++;
        if (hubs) {
          note(kWSigmaChunks,
               Run(g, "bc_sigma_chunks", 0, BcSigmaChunks, grid_c, blk, static_cast<const int32_t*>(g->in_dst),
                   static_cast<const int32_t*>(depth), sigma, d - 1, static_cast<const EdgeChunk*>(rchunks),
                   static_cast<const BcCounters*>(dc), wk(kWSigmaChunks)));
          Run(g, "bc_clear_reduce", 0, BcClearReduce, dim3(1), blk, dc);
          launches += 2;
        }
      }
      level = last;
      bucketed = true;
    } else {
      // Too deep for the buckets: the top-down discovery, level by level
      // (the BFS above only cost its own time; reset its outputs).
      Run(g, "bc_reset_source", 4 * N, BcResetSource, grid_n, blk, depth, n, roots[r], sigma, queue, dc);
      launches++;
    }
    while (!bucketed) {
      const int32_t qb = bounds[bounds.size() - 2], qn = bounds.back() - qb;
      const int64_t chunk_min = qn <= 4096 ? kExpandBS : kChunkEdges;
      const int idx = Run(g, "bc_discover", 8.0 * qn, BcDiscover, dim3(GridFor(qn, kBS, g->cu_count * 8)), blk,
                          static_cast<const int64_t*>(g->out_off), static_cast<const int32_t*>(g->out_dst),
                          depth, queue, qb, qn, level + 1, dc,
                          hubs ? chunks : static_cast<EdgeChunk*>(nullptr), chunk_min);
      launches++;
      if (hubs) {
        Run(g, "bc_discover_chunks", 0, BcDiscoverChunks, grid_c, blk, static_cast<const int32_t*>(g->out_dst),
            depth, queue, level + 1, dc, static_cast<const EdgeChunk*>(chunks));
        launches++;
      }
      const unsigned long long seq = MailNext(g);
      Run(g, "bc_publish", 0, BcPublish, dim3(1), blk, dc, g->mail_dev, seq);
      launches++;
      MailWait(g, seq, mail, kMailValues);
      CreditBytes(g, idx, 4.0 * mail[kMailExamined]);
      examined_total += static_cast<int64_t>(mail[kMailExamined]);
      const int32_t tail = static_cast<int32_t>(mail[kMailTail]);
      if (tail == bounds.back()) break;
      const int32_t nb = bounds.back(), nn = tail - nb;
      bounds.push_back(tail);
      note(kWSigma, Run(g, "bc_sigma", 20.0 * nn, BcSigma, dim3(GridFor(nn, kBS, g->cu_count * 16)), blk,
                        static_cast<const int64_t*>(g->in_off), static_cast<const int32_t*>(g->in_dst),
                        static_cast<const int32_t*>(depth), sigma, static_cast<const int32_t*>(queue), nb, nn,
                        level, rchunks, dc, wk(kWSigma)));
      launches++;
      if (hubs) {
        note(kWSigmaChunks,
             Run(g, "bc_sigma_chunks", 0, BcSigmaChunks, grid_c, blk, static_cast<const int32_t*>(g->in_dst),
                 static_cast<const int32_t*>(depth), sigma, level, static_cast<const EdgeChunk*>(rchunks),
                 static_cast<const BcCounters*>(dc), wk(kWSigmaChunks)));
        Run(g, "bc_clear_reduce", 0, BcClearReduce, dim3(1), blk, dc);
        launches += 2;
      }
      level++;
    }
    reached_total += bounds.back();
    levels_total += level + 1;
    // Backward: deepest level first; delta of level d reads level d + 1.
    for (int32_t d = level; d >= 0; d--) {
      const int32_t qb = bounds[d], qn = bounds[d + 1] - bounds[d];
      note(kWDelta, Run(g, "bc_delta", (16.0 + 2 * sizeof(T) + 8.0) * qn, BcDelta<T, kKahan>,
                        dim3(GridFor(qn, kBS, g->cu_count * 16)), blk, static_cast<const int64_t*>(g->out_off),
                        fold_dst, static_cast<const int32_t*>(depth), static_cast<const double*>(sigma), delta,
                        score, comp, static_cast<const int32_t*>(queue), qb, qn, d, wave_list, hub_list, dc,
                        wk(kWDelta)));
      launches++;
      if (g->max_degree >= kWaveEdges) {
        // Hubs fold on the side stream beside the wavefront kernel (disjoint
        // vertices; both only read level d + 1); both finish before the
        // list is cleared.
        const bool hubs = g->max_degree >= kBcHubEdges;
        if (hubs) {
          EnsureSideStream(g);
          GG_HIP(hipEventRecord(g->ev_fork, st));
          GG_HIP(hipStreamWaitEvent(g->side, g->ev_fork, 0));
          note(kWDeltaHub,
               RunOn(g, g->side, "bc_delta_hub", 0, BcDeltaHub<T, kKahan>, dim3(g->cu_count * kBcHubGrid), dim3(kBcHubBS),
                     static_cast<const int64_t*>(g->out_off), fold_dst, static_cast<const int32_t*>(depth),
                     static_cast<const double*>(sigma), delta, score, comp, static_cast<const int32_t*>(hub_list),
                     d, dc, wk(kWDeltaHub)));
          GG_HIP(hipEventRecord(g->ev_join, g->side));
          launches++;
        }
        note(kWDeltaWave,
             Run(g, "bc_delta_wave", 0, BcDeltaWave<T, kKahan>, dim3(GridFor(qn, kBS / kWave, g->cu_count * 8)),
                 blk, static_cast<const int64_t*>(g->out_off), fold_dst, static_cast<const int32_t*>(depth),
                 static_cast<const double*>(sigma), delta, score, comp, static_cast<const int32_t*>(wave_list), d,
                 static_cast<const BcCounters*>(dc), wk(kWDeltaWave)));
        if (hubs) GG_HIP(hipStreamWaitEvent(st, g->ev_join, 0));
        Run(g, "bc_clear_reduce", 0, BcClearReduce, dim3(1), blk, dc);
        launches += 2;
      }
    }
  }
  GG_HIP(hipMemsetAsync(max_bits, 0, sizeof(unsigned long long), st));
  Run(g, "bc_max", N * sizeof(T) * 2, BcMax<T>, dim3(GridFor(n, kBS, g->cu_count * 4)), blk,
      static_cast<const T*>(score), static_cast<const T*>(comp), n, max_bits);
  Run(g, "bc_normalize", N * (sizeof(T) * 2 + 4), BcNormalize<T>, grid_n, blk, static_cast<const T*>(score),
      static_cast<const T*>(comp), n, static_cast<const unsigned long long*>(max_bits), out32);
  launches += 2;
  GG_HIP(hipGetLastError());
  if (g->profile) {
    // Bytes beyond the per-vertex parts credited at launch: per list entry
    // its id and the neighbour's depth (8 B); per matched in-neighbour its
    // sigma (8 B); per successor its sigma and delta (8 + sizeof(T) B).
    const std::vector<unsigned long long> w = ReadWork(g, work, kWKinds);
    for (int kind = 0; kind < kWKinds; kind++) {
      const double per_match = kind <= kWSigmaChunks ? 8.0 : 8.0 + sizeof(T);
      CreditBytes(g, rec[kind], 8.0 * w[2 * kind] + per_match * w[2 * kind + 1]);
    }
  }
  FinishTiming(g, stats, wt);
  if (scores_out) CopyOutVertexF32(g, out32, scores_out);
  if (stats) {
    stats->total_ms = wt.ms();
    stats->iterations = levels_total;
    stats->reached = reached_total;
    stats->edges_examined = examined_total;
    stats->launches = launches;
  }
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_bc(gg_graph* g, const int32_t* roots, int k, int mode, float* scores, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "BC needs an unweighted graph upload");
    GG_REQUIRE(k >= 0 && (k == 0 || roots != nullptr), GG_EINVAL, "bad root list");
    GG_REQUIRE(mode == GG_BC_FP32 || mode == GG_BC_FP64 || mode == GG_BC_FP32_KAHAN, GG_EINVAL,
               "unknown precision mode");
    for (int i = 0; i < k; i++) GG_REQUIRE(roots[i] >= 0 && roots[i] < g->n, GG_EINVAL, "root out of range");
    GG_HIP(hipSetDevice(g->device));
    std::vector<int32_t> mapped(roots, roots + k);
    for (int32_t& r : mapped) r = DeviceVertex(g, r);
    roots = mapped.data();
    if (mode == GG_BC_FP32)
      RunBrandes<float, false>(g, roots, k, scores, stats, wt);
    else if (mode == GG_BC_FP32_KAHAN)
      RunBrandes<float, true>(g, roots, k, scores, stats, wt);
    else
      RunBrandes<double, false>(g, roots, k, scores, stats, wt);
  });
}
