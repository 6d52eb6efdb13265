// Batched Brandes betweenness centrality: up to 64 roots in one traversal.
//
// Reference: Brandes (src/bc.cc:95-150) runs one root at a time: a BFS
// that counts shortest paths (PBFS, bc.cc:51-92, path counts in double),
// then, level by level from the deepest to the root's own, the dependency
// delta_u = sum over successors v (depth(v) = depth(u) + 1, in u's sorted
// out-list order) of (sigma_u / sigma_v) * (1 + delta_v), accumulated in a
// float, and scores[u] += delta_u (bc.cc:121-137).  After all roots every
// score is divided by the largest (bc.cc:141-148).
//
// Batched design (vertex-major rows, lane = root):
//   * the batched BFS traversal (gg_msbfs.h) keeps every level's frontier
//     masks F_L (bit r of F_L(v): root r reached v at level L);
//   * sigma[v][r] (double, 64 per vertex): per level, every vertex with new
//     bits sums sigma_r over the in-neighbours u with bit r in F_{L-1}(u).
//     A wavefront serves one vertex, lane r handles root r, and the 64
//     sigma values of an in-neighbour are one 512-byte row read.  Path
//     counts are integers, so the order of these sums does not change them
//     (below 2^53); vertices with >= 512 in-edges are split into chunks
//     that add into the row with atomics;
//   * delta[u][r]: per level from the deepest, lane r folds the terms of
//     root r over u's out-list in order, with the reference's arithmetic
//     (float: the double term rounded into a float total; FP64 mode:
//     double throughout), so each delta_r(u) equals the single-root value;
//   * scores: per vertex, the deltas of the roots that reached it added in
//     root order (the order of the reference's trial loop), then the max
//     normalisation.  In FP32 mode the scores equal the reference's bit for
//     bit.
// Device memory: the frontier masks (8 bytes per vertex per level), sigma
// (512 bytes per vertex) and delta (256 or 512 bytes per vertex).
#include "gg_msbfs.h"

namespace gg {
namespace {

constexpr int kMbBS = 256;
constexpr int kSigmaLoads = 8;   // sigma rows in flight per wavefront step
constexpr int kDeltaLoads = 8;   // (sigma, delta) row pairs in flight per wavefront step
// Dependency folds of vertices with at least kHubFold successors run in
// MbDeltaHub: P producer wavefronts compute the terms of kHubEdges
// successors each per round into an LDS double buffer, one wavefront folds
// them in list order.  Most hubs go to small blocks (two producers, three
// wavefronts, 16 KB of LDS, at least seven wavefronts per SIMD: nine hubs on
// a CU at once); hubs of at least kHubWide successors, whose fold would
// otherwise outlast the rest of the level, go to wide blocks (fifteen
// producers, 120 successors per round).
constexpr int64_t kHubFold = 2048;
constexpr int64_t kHubWide = 65536;
constexpr int kHubEdges = 8;
constexpr int kHubSmall = 2;
constexpr int kHubLarge = 15;
constexpr int kHubSmallBlocksPerCu = 8;

__global__ void MbRootSigma(const int32_t* __restrict__ roots, int k, double* __restrict__ sigma) {
  const int r = threadIdx.x;
  if (r < k) sigma[int64_t(roots[r]) * 64 + r] = 1.0;
}

// Pops up to B set bits of `any` (lowest first) into j[]; the rest are -1.
// Wave-uniform.
template <int B>
__device__ __forceinline__ void PopBits(uint64_t& any, int (&j)[B]) {
#pragma unroll
  for (int t = 0; t < B; t++) {
    if (any) {
      j[t] = __builtin_ctzll(any);
      any &= any - 1;
    } else {
      j[t] = -1;
    }
  }
}

// Sigma of the vertices that gained bits at this level (cur = F_L, prev =
// F_{L-1}).  Wavefront w of W scans the vertices w, w + W, w + 2W, ... 64 at
// a time (consecutive ids on different wavefronts) and serves those with
// new bits one at a time.  Heavy vertices (>= kHeavyEdges in-edges)
// only get their entries zeroed here; MbSigmaHeavy adds their sums.
__global__ __launch_bounds__(kMbBS) void MbSigma(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                                 const uint64_t* __restrict__ cur, const uint64_t* __restrict__ prev,
                                                 double* __restrict__ sigma, int64_t n,
                                                 unsigned long long* work) {
  __shared__ unsigned long long s_red[kMbBS / kWave];
  const int lane = LaneId();
  const int64_t wave = (int64_t(blockIdx.x) * kMbBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kMbBS / kWave);
  unsigned long long ex = 0, pairs = 0;
  for (int64_t base = wave; base < n; base += nwaves * kWave) {
    const int64_t vl = base + lane * nwaves;
    uint64_t fl = 0;
    int64_t bl = 0, el = 0;
    if (vl < n) {
      fl = cur[vl];
      if (fl) {
        bl = in_off[vl];
        el = in_off[vl + 1];
      }
    }
    uint64_t act = __ballot(fl != 0);
    while (act) {
      const int i = __builtin_ctzll(act);
      act &= act - 1;
      const int64_t v = base + i * nwaves;
      const uint64_t fv = __shfl(fl, i);
      const int64_t b = __shfl(bl, i), e = __shfl(el, i);
      const bool mine = (fv >> lane) & 1;
      double* row = sigma + v * 64;
      if (e - b >= kHeavyEdges) {
        if (mine) row[lane] = 0.0;
        continue;
      }
      double acc = 0.0;
      for (int64_t cb = b; cb < e; cb += kWave) {
        const int64_t k = cb + lane;
        int32_t ul = 0;
        uint64_t ml = 0;
        if (k < e) {
          ul = in_dst[k];
          ml = prev[ul] & fv;
        }
        pairs += static_cast<unsigned long long>(__popcll(ml));
        uint64_t any = __ballot(ml != 0);
        while (any) {
          int j[kSigmaLoads];
          PopBits(any, j);
          double x[kSigmaLoads];
#pragma unroll
          for (int t = 0; t < kSigmaLoads; t++) {
            x[t] = 0.0;
            if (j[t] >= 0) {
              const uint64_t m = __shfl(ml, j[t]);
              const int32_t u = __shfl(ul, j[t]);
              if ((m >> lane) & 1) x[t] = sigma[int64_t(u) * 64 + lane];
            }
          }
#pragma unroll
          for (int t = 0; t < kSigmaLoads; t++) acc += x[t];
        }
      }
      ex += static_cast<unsigned long long>(e - b);
      if (mine) row[lane] = acc;
    }
  }
  ex = BlockSum<kMbBS>(lane == 0 ? ex : 0ull, s_red);
  pairs = BlockSum<kMbBS>(pairs, s_red);
  WorkAdd(work, ex, pairs);
}

// Sigma of heavy vertices: one wavefront per chunk of in-edges, partial
// sums added into the row (integers: exact in any order).
__global__ __launch_bounds__(kMbBS) void MbSigmaHeavy(const int32_t* __restrict__ in_dst,
                                                      const uint64_t* __restrict__ cur,
                                                      const uint64_t* __restrict__ prev, double* sigma,
                                                      const EdgeChunk* __restrict__ chunks, int32_t nch,
                                                      unsigned long long* work) {
  __shared__ unsigned long long s_red[kMbBS / kWave];
  const int lane = LaneId();
  const int64_t wave = (int64_t(blockIdx.x) * kMbBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kMbBS / kWave);
  unsigned long long ex = 0, pairs = 0;
  for (int64_t it = wave; it < nch; it += nwaves) {
    const EdgeChunk ch = chunks[it];
    const uint64_t fv = cur[ch.u];
    if (fv == 0) continue;
    double acc = 0.0;
    for (int64_t cb = ch.beg; cb < ch.end; cb += kWave) {
      const int64_t k = cb + lane;
      int32_t ul = 0;
      uint64_t ml = 0;
      if (k < ch.end) {
        ul = in_dst[k];
        ml = prev[ul] & fv;
      }
      pairs += static_cast<unsigned long long>(__popcll(ml));
      uint64_t any = __ballot(ml != 0);
      while (any) {
        int j[kSigmaLoads];
        PopBits(any, j);
        double x[kSigmaLoads];
#pragma unroll
        for (int t = 0; t < kSigmaLoads; t++) {
          x[t] = 0.0;
          if (j[t] >= 0) {
            const uint64_t m = __shfl(ml, j[t]);
            const int32_t u = __shfl(ul, j[t]);
            if ((m >> lane) & 1) x[t] = sigma[int64_t(u) * 64 + lane];
          }
        }
#pragma unroll
        for (int t = 0; t < kSigmaLoads; t++) acc += x[t];
      }
    }
    ex += static_cast<unsigned long long>(ch.end - ch.beg);
    if (acc != 0.0) unsafeAtomicAdd(&sigma[int64_t(ch.u) * 64 + lane], acc);
  }
  ex = BlockSum<kMbBS>(lane == 0 ? ex : 0ull, s_red);
  pairs = BlockSum<kMbBS>(pairs, s_red);
  WorkAdd(work, ex, pairs);
}

// The reference's dependency term and accumulation (float totals take each
// double term rounded in, bc.cc:128-131; `1 + deltas[v]` is a float sum).
__device__ __forceinline__ double Term(double su, double sv, float dv) {
  return __dmul_rn(__ddiv_rn(su, sv), static_cast<double>(__fadd_rn(1.0f, dv)));
}
__device__ __forceinline__ double Term(double su, double sv, double dv) {
  return __dmul_rn(__ddiv_rn(su, sv), __dadd_rn(1.0, dv));
}
__device__ __forceinline__ float Fold(float acc, double t) {
  return static_cast<float>(__dadd_rn(static_cast<double>(acc), t));
}
__device__ __forceinline__ double Fold(double acc, double t) { return __dadd_rn(acc, t); }

// Folds the dependencies of vertex u for the roots in fu (its bits at this
// level) over its out-list [b, e): lane r accumulates root r's terms in
// list order and stores delta_r(u).  Called by a whole wavefront; returns
// the entries read.  The next chunk's successors and masks are loaded
// before the current chunk's rows are folded, and up to kDeltaLoads row
// pairs are in flight per step.
template <class T>
__device__ __forceinline__ int64_t DeltaVertex(const int32_t* __restrict__ dst, const uint64_t* __restrict__ next,
                                               const double* __restrict__ sigma, T* __restrict__ delta, int64_t u,
                                               uint64_t fu, int64_t b, int64_t e, unsigned long long& pairs) {
  const int lane = LaneId();
  const bool mine = (fu >> lane) & 1;
  const double su = mine ? sigma[u * 64 + lane] : 1.0;
  T acc = T(0);
  int32_t vn = 0;
  uint64_t mn = 0;
  if (b + lane < e) {
    vn = dst[b + lane];
    mn = next[vn] & fu;
  }
  for (int64_t cb = b; cb < e; cb += kWave) {
    const int32_t vl = vn;
    const uint64_t ml = mn;
    pairs += static_cast<unsigned long long>(__popcll(ml));
    const int64_t kn = cb + kWave + lane;
    vn = 0;
    mn = 0;
    if (kn < e) {
      vn = dst[kn];
      mn = next[vn] & fu;
    }
    uint64_t any = __ballot(ml != 0);
    while (any) {
      int j[kDeltaLoads];
      PopBits(any, j);
      double sv[kDeltaLoads];
      T dv[kDeltaLoads];
      bool on[kDeltaLoads];
#pragma unroll
      for (int t = 0; t < kDeltaLoads; t++) {
        on[t] = false;
        sv[t] = 1.0;
        dv[t] = T(0);
        if (j[t] >= 0) {
          const uint64_t m = __shfl(ml, j[t]);
          const int32_t v = __shfl(vl, j[t]);
          on[t] = (m >> lane) & 1;
          if (on[t]) {
            sv[t] = sigma[int64_t(v) * 64 + lane];
            dv[t] = delta[int64_t(v) * 64 + lane];
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kDeltaLoads; t++)
        if (on[t]) acc = Fold(acc, Term(su, sv[t], dv[t]));
    }
  }
  if (mine) delta[u * 64 + lane] = acc;
  return e - b;
}

// Dependencies of the vertices at level L (cur = F_L, next = F_{L+1}, null
// at the deepest level).  Wavefront w of W scans the vertices w, w + W,
// w + 2W, ... 64 at a time and serves those with bits at L one at a time.
// The fold of each (vertex, root) is sequential by definition, so a hub's
// wavefront is the longest one; the stride puts consecutive ids (in a
// degree-ordered layout, hubs of similar size) on different wavefronts.
template <class T>
__global__ __launch_bounds__(kMbBS) void MbDelta(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                 const uint64_t* __restrict__ cur, const uint64_t* __restrict__ next,
                                                 const double* __restrict__ sigma, T* __restrict__ delta, int64_t n,
                                                 unsigned long long* work) {
  __shared__ unsigned long long s_red[kMbBS / kWave];
  const int lane = LaneId();
  const int64_t wave = (int64_t(blockIdx.x) * kMbBS + threadIdx.x) / kWave;
  const int64_t nwaves = int64_t(gridDim.x) * (kMbBS / kWave);
  unsigned long long ex = 0, pairs = 0;
  for (int64_t base = wave; base < n; base += nwaves * kWave) {
    const int64_t ul = base + lane * nwaves;
    uint64_t fl = 0;
    int64_t bl = 0, el = 0;
    if (ul < n) {
      fl = cur[ul];
      if (fl) {
        bl = off[ul];
        el = off[ul + 1];
        if (el - bl >= kHubFold) fl = 0;     // MbDeltaHub's
        if (!next) el = bl;                   // deepest level: no successors
      }
    }
    uint64_t act = __ballot(fl != 0);
    while (act) {
      const int i = __builtin_ctzll(act);
      act &= act - 1;
      ex += static_cast<unsigned long long>(
          DeltaVertex<T>(dst, next, sigma, delta, base + i * nwaves, __shfl(fl, i), __shfl(bl, i), __shfl(el, i),
                         pairs));
    }
  }
  ex = BlockSum<kMbBS>(lane == 0 ? ex : 0ull, s_red);
  pairs = BlockSum<kMbBS>(pairs, s_red);
  WorkAdd(work, ex, pairs);
}

// Dependencies of the hubs (>= kHubFold successors) at level L, one block
// per hub.  Per round, producer wavefront p computes, for the successors
// b + round * kHubRound + p * kHubEdges + j (j < kHubEdges), lane r's term
// for root r (or -1 where the successor is not at level L + 1 for r) into
// one half of an LDS double buffer, while the consumer wavefront folds the
// previous round's half in successor order -- the same terms, in the same
// order, with the same arithmetic as DeltaVertex, so the result is
// bit-identical; only the loads and divisions run in parallel.
//
// A producer loads successor ids two rounds and their level masks one round
// before use, so a round waits on one load, its sigma/delta rows.  Row loads
// are unconditional (a lane whose successor is not on its root's next level
// reads the hub's own row and its term is dropped), which keeps the round's
// code free of branches and its one wait exact.
template <class T, int P>
__global__ __launch_bounds__((P + 1) * kWave, P <= kHubSmall ? 7 : 1) void MbDeltaHub(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                     const uint64_t* __restrict__ cur,
                                                     const uint64_t* __restrict__ next,
                                                     const double* __restrict__ sigma, T* __restrict__ delta,
                                                     const int32_t* __restrict__ hubs, int32_t nhubs,
                                                     unsigned long long* work) {
  constexpr int kHubRound = P * kHubEdges;
  __shared__ double terms[2][kHubRound * kWave];
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  unsigned long long ex = 0, pairs = 0;
  for (int32_t h = blockIdx.x; h < nhubs; h += gridDim.x) {
    const int64_t u = hubs[h];
    const uint64_t fu = cur[u];
    if (fu == 0) continue;  // uniform per block
    const int64_t b = off[u];
    const int64_t e = next ? off[u + 1] : b;  // deepest level: no successors
    const int64_t rounds = (e - b + kHubRound - 1) / kHubRound;
    if (wid < P) {
      const double su = ((fu >> lane) & 1) ? sigma[u * 64 + lane] : 1.0;
      const int64_t mine = b + wid * kHubEdges;  // slot 0 of round 0
      // Lane j < kHubEdges: successor slot j of round k (any in-range entry
      // past the end; such slots are dropped below).
      auto id_at = [&](int64_t k) -> int32_t {
        const int64_t i = mine + k * kHubRound + lane;
        return dst[lane < kHubEdges && i < e ? i : b];
      };
      int32_t v_cur = 0, v_nxt = 0;
      uint64_t m_cur = 0;
      if (rounds > 0) {
        v_cur = id_at(0);
        v_nxt = id_at(1);
        m_cur = next[v_cur];
      }
      for (int64_t k = 0; k < rounds; k++) {
        const int32_t v_nxt2 = id_at(k + 2);
        const uint64_t m_nxt = next[v_nxt];
        const int64_t first = mine + k * kHubRound;
        double sv[kHubEdges];
        T dv[kHubEdges];
        uint32_t on = 0;
#pragma unroll
        for (int j = 0; j < kHubEdges; j++) {
          const int32_t v = __shfl(v_cur, j);
          const uint64_t m = __shfl(m_cur, j) & fu;
          const bool o = first + j < e && ((m >> lane) & 1);
          const int64_t row = (o ? int64_t(v) : u) * 64 + lane;
          sv[j] = sigma[row];
          dv[j] = delta[row];
          on |= static_cast<uint32_t>(o) << j;
        }
        pairs += static_cast<unsigned long long>(__popc(on));
        double* out = terms[k & 1] + wid * kHubEdges * kWave;
#pragma unroll
        for (int j = 0; j < kHubEdges; j++) {
          const double t = Term(su, sv[j], dv[j]);
          out[j * kWave + lane] = ((on >> j) & 1) ? t : -1.0;
        }
        v_cur = v_nxt;
        v_nxt = v_nxt2;
        m_cur = m_nxt;
        __syncthreads();
      }
      __syncthreads();  // the consumer's last fold
    } else {
      T acc = T(0);
      auto fold = [&](int64_t k) {
        const double* in = terms[k & 1];
        const int64_t first = b + k * kHubRound;
        const int cnt = static_cast<int>(e - first < kHubRound ? e - first : kHubRound);
        int j = 0;
        for (; j + 8 <= cnt; j += 8) {  // eight LDS reads in flight, then the ordered fold
          double t[8];
#pragma unroll
          for (int q = 0; q < 8; q++) t[q] = in[(j + q) * kWave + lane];
#pragma unroll
          for (int q = 0; q < 8; q++)
            if (t[q] >= 0.0) acc = Fold(acc, t[q]);
        }
        for (; j < cnt; j++) {
          const double t = in[j * kWave + lane];
          if (t >= 0.0) acc = Fold(acc, t);
        }
      };
      for (int64_t k = 0; k < rounds; k++) {
        if (k >= 1) fold(k - 1);
        __syncthreads();
      }
      if (rounds > 0) fold(rounds - 1);
      __syncthreads();
      if ((fu >> lane) & 1) delta[u * 64 + lane] = acc;
      if (lane == 0) ex += static_cast<unsigned long long>(e - b);
    }
  }
  const int s = blockIdx.x & (kWorkStripes - 1);
  if (wid == P && lane == 0 && ex) atomicAdd(&work[s], ex);
  if (wid < P) {
    pairs = WaveSum(pairs);
    if (lane == 0 && pairs) atomicAdd(&work[kWorkStripes + s], pairs);
  }
}

// Vertices with at least kHubFold out-edges, split at kHubWide (once per
// graph).
__global__ __launch_bounds__(kMbBS) void MbHubList(const int64_t* __restrict__ off, int64_t n, int32_t* small,
                                                   int32_t* wide, int32_t* tails) {
  for (int64_t base = int64_t(blockIdx.x) * kMbBS; base < n; base += int64_t(gridDim.x) * kMbBS) {
    const int64_t v = base + threadIdx.x;
    const int64_t d = v < n ? off[v + 1] - off[v] : 0;
    WaveAppend(d >= kHubFold && d < kHubWide, static_cast<int32_t>(v), small, tails);
    WaveAppend(d >= kHubWide, static_cast<int32_t>(v), wide, tails + 1);
  }
}

__device__ __forceinline__ float AddT(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ double AddT(double a, double b) { return __dadd_rn(a, b); }

// score[u] = the deltas of the roots that reached u, added in root order.
// delta holds one 64-entry row per vertex, so a thread summing its own row
// would touch a different cache line in every lane of a load. Instead each
// wave takes 64 vertices: it reads their rows one row per load (lane r reads
// root r, only where root r reached that vertex) into an LDS tile whose rows
// are padded to 65 entries, then lane j adds row j from LDS in root order,
// the order of the one-thread loop, so the sums are unchanged bit for bit.
constexpr int kMbScoreRow = 65;
constexpr int kMbScoreBatch = 32;  // rows loaded before they are written to LDS
template <class T>
__global__ __launch_bounds__(kMbBS) void MbScores(const uint64_t* __restrict__ seen, const T* __restrict__ delta,
                                                  int64_t n, T* __restrict__ score) {
  __shared__ T tile[kMbBS / kWave][kWave * kMbScoreRow];
  const int lane = threadIdx.x & (kWave - 1);
  T* row = tile[threadIdx.x / kWave];
  for (int64_t base = blockIdx.x * int64_t(kMbBS); base < n; base += int64_t(gridDim.x) * kMbBS) {
    const int64_t v0 = base + (threadIdx.x - lane);
    const int64_t u = v0 + lane;
    const uint64_t mine = u < n ? seen[u] : 0;
    const uint32_t lo = static_cast<uint32_t>(mine), hi = static_cast<uint32_t>(mine >> 32);
    for (int i0 = 0; i0 < kWave; i0 += kMbScoreBatch) {
      T v[kMbScoreBatch];
#pragma unroll
      for (int i = 0; i < kMbScoreBatch; ++i) {
        // Bit `lane` of seen[v0 + i0 + i]; rows past n have seen 0.
        const uint32_t wlo = __builtin_amdgcn_readlane(lo, i0 + i), whi = __builtin_amdgcn_readlane(hi, i0 + i);
        const bool on = ((lane < 32 ? wlo >> lane : whi >> (lane - 32)) & 1u) != 0;
        v[i] = on ? delta[(v0 + i0 + i) * kWave + lane] : T(0);
      }
#pragma unroll
      for (int i = 0; i < kMbScoreBatch; ++i) row[(i0 + i) * kMbScoreRow + lane] = v[i];
    }
    __syncthreads();
    T acc = T(0);
    uint64_t s = mine;
    while (s) {
      const int r = __builtin_ctzll(s);
      s &= s - 1;
      acc = AddT(acc, row[lane * kMbScoreRow + r]);
    }
    if (u < n) score[u] = acc;
    __syncthreads();  // the next pass overwrites the tile
  }
}

// Largest score (scores are >= 0, so their bit patterns order like the
// values), then every score divided by it (bc.cc:141-148).
template <class T>
__global__ __launch_bounds__(kMbBS) void MbMax(const T* __restrict__ score, int64_t n, unsigned long long* max_bits) {
  __shared__ double s_max[kMbBS / kWave];
  double m = 0.0;
  for (int64_t i = blockIdx.x * int64_t(kMbBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kMbBS) {
    const double v = static_cast<double>(score[i]);
    m = v > m ? v : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(m, o);
    m = t > m ? t : m;
  }
  if (LaneId() == 0) s_max[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kMbBS / kWave; w++) m = s_max[w] > m ? s_max[w] : m;
    atomicMax(max_bits, static_cast<unsigned long long>(__double_as_longlong(m)));
  }
}

template <class T>
__global__ __launch_bounds__(kMbBS) void MbNormalize(const T* __restrict__ score, int64_t n,
                                                     const unsigned long long* max_bits, float* __restrict__ out) {
  const double mx = __longlong_as_double(static_cast<long long>(*max_bits));
  const T biggest = static_cast<T>(mx);
  for (int64_t i = blockIdx.x * int64_t(kMbBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kMbBS)
    out[i] = static_cast<float>(score[i] / biggest);
}

__global__ void MbClear(unsigned long long* p, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) p[i] = 0;
}

template <class T>
void RunBatchBrandes(gg_graph* g, const int32_t* droots, int k, float* scores, gg_stats* stats,
                     const WallTimer& wt) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  const dim3 blk(kMbBS);
  const dim3 grid_n(GridFor(n, kMbBS, g->cu_count * 16));
  const dim3 grid_w(GridFor(n, kMbBS, g->cu_count * 16));  // each wavefront takes 64 vertices at a time
  const char* tag = sizeof(T) == 4 ? "mb.delta32" : "mb.delta64";
  double* sigma = g->Scratch<double>("mb.sigma", n * 64);
  T* delta = g->Scratch<T>(tag, n * 64);
  T* score = g->Scratch<T>(sizeof(T) == 4 ? "mb.score32" : "mb.score64", n);
  float* out = g->Scratch<float>("mb.out", n);
  // Work counters per kernel kind (list entries read, matched (entry, root)
  // pairs), then the max-score bits.
  enum { kWSigma, kWSigmaHeavy, kWDelta, kWHub, kWHubWide, kWKinds };
  constexpr int kCtr = kWKinds * kWorkCounters + 1;
  unsigned long long* ctr = g->Scratch<unsigned long long>("mb.ctr", kCtr);
  unsigned long long* max_bits = ctr + kWKinds * kWorkCounters;
  auto wk = [&](int kind) { return ctr + kind * kWorkCounters; };
  int rec[kWKinds] = {-1, -1, -1, -1, -1};
  auto note = [&](int kind, int idx) {
    if (rec[kind] < 0) rec[kind] = idx;
  };
  const double N = static_cast<double>(n);

  // Dependencies are folded in each vertex's uploaded list order (the
  // reference's successor order), also on a degree-ordered handle.
  const int32_t* fold_dst = UploadOrderLists(g);
  BeginTiming(g);
  MsTraversal t = MsTraverse(g, droots, nullptr, k, k == 64 ? ~0ull : (1ull << k) - 1, 15, nullptr, true);
  const int levels = static_cast<int>(t.frontiers.size());
  int64_t launches = t.launches;
  Run(g, "msbc_clear", 0, MbClear, dim3(1), blk, ctr, kCtr);
  Run(g, "msbc_root_sigma", 0, MbRootSigma, dim3(1), dim3(64), droots, k, sigma);
  launches += 2;
  const int32_t nhchunks = static_cast<int32_t>(g->ms_heavy_chunks);
  const EdgeChunk* hchunks = g->Scratch<EdgeChunk>("ms.hchunks", 1);
  for (int L = 1; L < levels; L++) {
    note(kWSigma, Run(g, "msbc_sigma", 16.0 * N, MbSigma, grid_w, blk, static_cast<const int64_t*>(g->in_off),
                      static_cast<const int32_t*>(g->in_dst), static_cast<const uint64_t*>(t.frontiers[L]),
                      static_cast<const uint64_t*>(t.frontiers[L - 1]), sigma, n, wk(kWSigma)));
    note(kWSigmaHeavy,
         Run(g, "msbc_sigma_heavy", 0, MbSigmaHeavy,
             dim3(GridFor(int64_t(nhchunks) * kWave, kMbBS, g->cu_count * 16)), blk,
             static_cast<const int32_t*>(g->in_dst), static_cast<const uint64_t*>(t.frontiers[L]),
             static_cast<const uint64_t*>(t.frontiers[L - 1]), sigma, hchunks, nhchunks, wk(kWSigmaHeavy)));
    launches += 2;
  }
  // Hubs (their folds split across a block, MbDeltaHub): listed once per
  // graph, at most m / kHubFold of them.
  const int64_t max_hubs = std::min<int64_t>(n, g->out_entries / kHubFold + 1);
  int32_t* hubs = g->Scratch<int32_t>("mb.hubs", max_hubs);
  int32_t* wide = g->Scratch<int32_t>("mb.wide_hubs", std::min<int64_t>(n, g->out_entries / kHubWide + 1));
  if (g->mb_hubs < 0) {
    int32_t* tails = g->Scratch<int32_t>("mb.hub_tails", 2);
    GG_HIP(hipMemsetAsync(tails, 0, 2 * sizeof(int32_t), st));
    Run(g, "msbc_hub_list", 0, MbHubList, dim3(GridFor(n, kMbBS, g->cu_count * 4)), blk,
        static_cast<const int64_t*>(g->out_off), n, hubs, wide, tails);
    int32_t h[2] = {0, 0};
    GG_HIP(hipMemcpyAsync(h, tails, sizeof(h), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    g->mb_hubs = h[0];
    g->mb_big_hubs = h[1];
  }
  const int32_t nhubs = static_cast<int32_t>(g->mb_hubs), nwide = static_cast<int32_t>(g->mb_big_hubs);
  if (nhubs + nwide > 0) EnsureSideStream(g);
  for (int L = levels - 1; L >= 0; L--) {
    const uint64_t* next = L + 1 < levels ? t.frontiers[L + 1] : nullptr;
    // The hubs' folds (side streams, wide blocks first) overlap the other
    // vertices' (main stream): they write disjoint rows and read only level
    // L + 1's.
    if (nhubs + nwide > 0) GG_HIP(hipEventRecord(g->ev_fork, st));
    if (nwide > 0) {
      GG_HIP(hipStreamWaitEvent(g->side, g->ev_fork, 0));
      note(kWHubWide, RunOn(g, g->side, "msbc_delta_hub_wide", 0, MbDeltaHub<T, kHubLarge>,
                            dim3(GridFor(nwide, 1, g->cu_count)), dim3((kHubLarge + 1) * kWave),
                            static_cast<const int64_t*>(g->out_off), fold_dst,
                            static_cast<const uint64_t*>(t.frontiers[L]), next, static_cast<const double*>(sigma),
                            delta, static_cast<const int32_t*>(wide), nwide, wk(kWHubWide)));
      GG_HIP(hipEventRecord(g->ev_join, g->side));
      launches++;
    }
    if (nhubs > 0) {
      GG_HIP(hipStreamWaitEvent(g->side2, g->ev_fork, 0));
      note(kWHub, RunOn(g, g->side2, "msbc_delta_hub", 0, MbDeltaHub<T, kHubSmall>,
                        dim3(GridFor(nhubs, 1, g->cu_count * kHubSmallBlocksPerCu)), dim3((kHubSmall + 1) * kWave),
                        static_cast<const int64_t*>(g->out_off), fold_dst,
                        static_cast<const uint64_t*>(t.frontiers[L]), next, static_cast<const double*>(sigma), delta,
                        static_cast<const int32_t*>(hubs), nhubs, wk(kWHub)));
      GG_HIP(hipEventRecord(g->ev_join2, g->side2));
      launches++;
    }
    note(kWDelta, Run(g, "msbc_delta", 16.0 * N, MbDelta<T>, grid_w, blk, static_cast<const int64_t*>(g->out_off),
                      fold_dst, static_cast<const uint64_t*>(t.frontiers[L]), next,
                      static_cast<const double*>(sigma), delta, n, wk(kWDelta)));
    launches++;
    if (nwide > 0) GG_HIP(hipStreamWaitEvent(st, g->ev_join, 0));
    if (nhubs > 0) GG_HIP(hipStreamWaitEvent(st, g->ev_join2, 0));
  }
  Run(g, "msbc_scores", 8.0 * N + sizeof(T) * N, MbScores<T>, grid_n, blk, static_cast<const uint64_t*>(t.seen),
      static_cast<const T*>(delta), n, score);
  Run(g, "msbc_max", sizeof(T) * N, MbMax<T>, grid_n, blk, static_cast<const T*>(score), n, max_bits);
  Run(g, "msbc_normalize", (sizeof(T) + 4.0) * N, MbNormalize<T>, grid_n, blk, static_cast<const T*>(score), n,
      static_cast<const unsigned long long*>(max_bits), out);
  launches += 3;
  GG_HIP(hipGetLastError());
  if (g->profile) {
    // Bytes beyond the 16 B per vertex credited at launch: per list entry
    // its id and the neighbour's level mask (12 B); per matched (entry,
    // root) pair the neighbour's sigma (8 B) when pulling sigma, its sigma
    // and delta (8 + sizeof(T) B) when folding dependencies.
    const std::vector<unsigned long long> w = ReadWork(g, ctr, kWKinds);
    for (int kind = 0; kind < kWKinds; kind++) {
      const double per_pair = kind <= kWSigmaHeavy ? 8.0 : 8.0 + sizeof(T);
      CreditBytes(g, rec[kind], 12.0 * w[2 * kind] + per_pair * w[2 * kind + 1]);
    }
  }
  FinishTiming(g, stats, wt);
  if (scores) CopyOutVertexF32(g, out, scores);
  if (stats) {
    const std::vector<unsigned long long> w = ReadWork(g, ctr, kWKinds);
    unsigned long long* sums = g->Scratch<unsigned long long>("ms.reach", 2);
    MsZero2<<<1, 64, 0, st>>>(sums);
    MsReach<<<grid_n, dim3(kBS), 0, st>>>(t.seen, g->out_off, n, sums);
    unsigned long long reach[2];
    GG_HIP(hipMemcpyAsync(reach, sums, sizeof(reach), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
    unsigned long long ex = 0;
    for (int kind = 0; kind < kWKinds; kind++) ex += w[2 * kind];
    stats->total_ms = wt.ms();
    stats->iterations = levels;
    stats->td_steps = t.td;
    stats->bu_steps = t.bu;
    stats->reached = static_cast<int64_t>(reach[0]);
    stats->reached_edges = static_cast<int64_t>(reach[1]);
    stats->edges_examined = t.examined + static_cast<int64_t>(ex);
    stats->launches = launches;
  }
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_bc_batch(gg_graph* g, const int32_t* roots, int k, int mode, float* scores, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "BC needs an unweighted graph upload");
    GG_REQUIRE(k >= 1 && k <= 64 && roots, GG_EINVAL, "1 to 64 roots");
    GG_REQUIRE(mode == GG_BC_FP32 || mode == GG_BC_FP64, mode == GG_BC_FP32_KAHAN ? GG_ENOTSUP : GG_EINVAL,
               "batched BC runs in GG_BC_FP32 or GG_BC_FP64");
    for (int i = 0; i < k; i++) GG_REQUIRE(roots[i] >= 0 && roots[i] < g->n, GG_EINVAL, "root out of range");
    GG_HIP(hipSetDevice(g->device));
    int32_t* droots = g->Scratch<int32_t>("ms.roots", 64);
    GG_HIP(hipMemcpyAsync(droots, roots, sizeof(int32_t) * k, hipMemcpyHostToDevice, g->stream));
    DeviceVertices(g, droots, k);
    if (mode == GG_BC_FP32)
      RunBatchBrandes<float>(g, droots, k, scores, stats, wt);
    else
      RunBatchBrandes<double>(g, droots, k, scores, stats, wt);
  });
}
