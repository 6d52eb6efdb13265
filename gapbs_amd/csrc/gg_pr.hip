// Pull-direction PageRank on gfx950.
//
// Reference: PageRankPull (src/pr_spmv.cc:35-61), the Jacobi form: every
// iteration first sets contrib[n] = score[n] / out_degree(n) for all n, then
// score[u] = (1 - d) / N + d * sum_{v in in(u)} contrib[v], with the L1
// change accumulated in double and the loop stopped when it drops below
// epsilon or after max_iters (pr_spmv.cc:41-58).  d = 0.85f and both
// constants are computed in float as the reference does (pr_spmv.cc:32,
// 38-39); there is no redistribution of dangling mass.
//
// GPU design: vertices are binned once per graph by in-degree -- below 32
// one thread per vertex, below 1024 one wavefront, else one block -- and an
// iteration is one gather launch over the three bins.  Each gather writes the
// vertex's new score and, for the next iteration, its new contribution into a
// second buffer, so the reference's separate contrib pass is fused away.
//
// Jacobi layouts.  On graphs of 2^20 to 2^27 vertices the iterations run
// through propagation blocks (section "Propagation blocks" below): two
// streaming phases per iteration, no random reads, so the bytes moved scale
// with the element size.  Source segments (next paragraph) are the layout
// when a segment width is set explicitly or blocks are turned off.
//
// Source segments (Jacobi modes).  A pull gather reads contrib[v] for every
// in-entry at random.  Measured on the MI355X, random 4-byte gathers run at
// about 1350 G/s from LDS, 200 G/s from an L2-resident range and 52 G/s from
// HBM: on a large graph most pull gathers miss L2 and move a whole cache
// line for 4 or 8 useful bytes.  With sorted in-lists, the entries of u
// whose source lies in one range of sources (a segment) form one contiguous
// run, a (destination, sub-list) pair of that segment.  Built once per graph
// and precision: the pairs of the leading segments that hold a few entries
// per pair, grouped by segment, with their sub-lists copied into one
// segment-major array.  A segment covers 2 MB of contributions (2^19 float
// or 2^18 double sources), so they stay in every XCD's L2 during its launch.
// An iteration runs one launch per segment, each adding its pairs' partial
// sums into the destinations' accumulators, then the usual three-bin gather
// over the rest of every list (the sources beyond the last segment), which
// starts from the accumulator and finishes the vertex.  On a degree-ordered
// Kronecker graph the hubs sit at the low ids, so a few segments hold most
// entries (scale 26: 16 segments hold 95% of them in 1.8 pairs per vertex).
// Optional LDS segments (gg_tuning.pr_lds_log2) come first and stage 128 KB
// of contributions in LDS per launch, one 1024-thread block per CU; at
// Kronecker 26 they measured slower than L2 segments (their per-pair
// gathers are latency-bound at 16 waves per CU), so they are off by default.
// Two precisions:
//   GG_PR_FP32  float scores and contributions; in the gathers each vertex's
//               sum is kept as a compensated (Neumaier) pair and merged
//               across lanes with TwoSum, so the sum carries about twice
//               float's precision; through propagation blocks it is exact
//               (64-bit fixed point), rounded to float once;
//   GG_PR_FP64  double scores, contributions and sums.
// The convergence residual is always accumulated in double.
#include "gg_common.h"
#include "gg_sort.h"

#include <algorithm>

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kStripes = 64;
constexpr int kGsPasses = 4;  // Gauss-Seidel sweep: vertex-id ranges updated in turn

// Compensated float sum: value = s + c.
struct CompSum {
  float s, c;
  __device__ __forceinline__ static CompSum Zero() { return CompSum{0.0f, 0.0f}; }
  __device__ __forceinline__ void Add(float x) {
    const float t = __fadd_rn(s, x);
    const float e = fabsf(s) >= fabsf(x) ? __fadd_rn(__fsub_rn(s, t), x) : __fadd_rn(__fsub_rn(x, t), s);
    c = __fadd_rn(c, e);
    s = t;
  }
  __device__ __forceinline__ void Merge(const CompSum& o) {
    const float t = __fadd_rn(s, o.s);
    const float bp = __fsub_rn(t, s);
    const float e = __fadd_rn(__fsub_rn(s, __fsub_rn(t, bp)), __fsub_rn(o.s, bp));  // TwoSum error
    c = __fadd_rn(__fadd_rn(c, o.c), e);
    s = t;
  }
  __device__ __forceinline__ float Value() const { return __fadd_rn(s, c); }
};

struct PlainSum {
  double s;
  __device__ __forceinline__ static PlainSum Zero() { return PlainSum{0.0}; }
  __device__ __forceinline__ void Add(double x) { s = __dadd_rn(s, x); }
  __device__ __forceinline__ void Merge(const PlainSum& o) { s = __dadd_rn(s, o.s); }
  __device__ __forceinline__ double Value() const { return s; }
};

// Accumulator type of each precision, and its form in memory (the
// per-vertex partial sums the segment launches build up).
template <class T>
struct Acc;
template <>
struct Acc<float> {
  using Type = CompSum;
  using Stored = float2;
};
template <>
struct Acc<double> {
  using Type = PlainSum;
  using Stored = double;
};

__device__ __forceinline__ CompSum Unpack(float2 v) { return CompSum{v.x, v.y}; }
__device__ __forceinline__ PlainSum Unpack(double v) { return PlainSum{v}; }
__device__ __forceinline__ float2 Pack(const CompSum& a) { return make_float2(a.s, a.c); }
__device__ __forceinline__ double Pack(const PlainSum& a) { return a.s; }

// Reads vertex u's accumulated partial sum and clears it for the next
// iteration.
template <class S>
__device__ __forceinline__ auto TakeAcc(S* acc, int32_t u) {
  const S v = acc[u];
  acc[u] = S{};
  return Unpack(v);
}

template <class A>
__device__ __forceinline__ A ShflXor(const A& a, int o);
template <>
__device__ __forceinline__ CompSum ShflXor(const CompSum& a, int o) {
  CompSum r = CompSum::Zero();
  r.s = __shfl_xor(a.s, o);
  r.c = __shfl_xor(a.c, o);
  return r;
}
template <>
__device__ __forceinline__ PlainSum ShflXor(const PlainSum& a, int o) {
  PlainSum r = PlainSum::Zero();
  r.s = __shfl_xor(a.s, o);
  return r;
}

template <class A>
__device__ __forceinline__ A WaveMerge(A a) {
  for (int o = 32; o > 0; o >>= 1) a.Merge(ShflXor(a, o));
  return a;
}

// score = base + d * sum, in the reference's precision and rounding (a
// product then a sum, not a fused multiply-add).
__device__ __forceinline__ float Combine(float base, float d, float sum) {
  return __fadd_rn(base, __fmul_rn(d, sum));
}
__device__ __forceinline__ double Combine(double base, double d, double sum) {
  return __dadd_rn(base, __dmul_rn(d, sum));
}

template <class T>
struct PrArgs {
  const int64_t* in_off;
  const int32_t* in_dst;
  const int64_t* out_off;
  const T* contrib_in;
  T* contrib_out;
  T* score;
  T base, damp;
  double* residual;  // striped
  int32_t lo, hi;    // vertex ids updated by this launch: [lo, hi)
  // With source segments: each list is gathered from in_off[u] + tail[u]
  // (the entries the segment launches did not take) and its sum starts from
  // acc[u], which is cleared.  Both null: whole lists, sums from zero.
  const int32_t* tail;
  typename Acc<T>::Stored* acc;
};

// The gather's starting point and initial sum for vertex u.
template <class T>
__device__ __forceinline__ int64_t ListBegin(const PrArgs<T>& a, int32_t u) {
  return a.in_off[u] + (a.tail ? a.tail[u] : 0);
}
template <class T>
__device__ __forceinline__ typename Acc<T>::Type StartSum(const PrArgs<T>& a, int32_t u, bool take) {
  using A = typename Acc<T>::Type;
  return a.acc && take ? TakeAcc(a.acc, u) : A::Zero();
}

// Adds contrib[ids[k]] for k = k0, k0 + step, ... < e to acc, with four
// independent gathers in flight per lane.
template <class T, class A>
__device__ __forceinline__ void GatherIds(const T* __restrict__ contrib, const int32_t* __restrict__ ids, A& acc,
                                          int64_t k0, int64_t e, int step) {
  int64_t k = k0;
  for (; k + 3 * step < e; k += 4 * step) {
    const int32_t v0 = ids[k], v1 = ids[k + step], v2 = ids[k + 2 * step], v3 = ids[k + 3 * step];
    const T x0 = contrib[v0], x1 = contrib[v1], x2 = contrib[v2], x3 = contrib[v3];
    acc.Add(x0);
    acc.Add(x1);
    acc.Add(x2);
    acc.Add(x3);
  }
  for (; k < e; k += step) acc.Add(contrib[ids[k]]);
}

template <class T, class A>
__device__ __forceinline__ void GatherStrided(const PrArgs<T>& a, A& acc, int64_t k0, int64_t e, int step) {
  GatherIds(a.contrib_in, a.in_dst, acc, k0, e, step);
}

// Finishes vertex u: new score, next contribution, |change|.
template <class T>
__device__ __forceinline__ double Finish(const PrArgs<T>& a, int32_t u, T sum) {
  const T old = a.score[u];
  const T now = Combine(a.base, a.damp, sum);
  a.score[u] = now;
  const int64_t deg = a.out_off[u + 1] - a.out_off[u];
  a.contrib_out[u] = now / static_cast<T>(deg);
  return fabs(static_cast<double>(now) - static_cast<double>(old));
}

__device__ __forceinline__ void AddResidual(double* residual, double r, double* s_red) {
  r = BlockSum<kBS>(r, s_red);
  if (threadIdx.x == 0 && r != 0.0) atomicAdd(&residual[blockIdx.x & (kStripes - 1)], r);
}

// The three gathers as parts of one launch: blocks [0, nb_large) serve the
// large bin (one block per vertex), the next nb_medium blocks the medium
// bin (one wavefront per vertex), the rest the small bin (one thread per
// vertex); `blk` and `nblk` are the block's index and count within its part.
template <class T>
__device__ __forceinline__ double GatherThreadPart(const PrArgs<T>& a, const int32_t* __restrict__ verts,
                                                  int32_t nv, int64_t blk, int64_t nblk) {
  using A = typename Acc<T>::Type;
  double r = 0.0;
  for (int64_t i = blk * kBS + threadIdx.x; i < nv; i += nblk * kBS) {
    const int32_t u = verts[i];
    if (u < a.lo || u >= a.hi) continue;
    const int64_t b = ListBegin(a, u), e = a.in_off[u + 1];
    A acc = StartSum(a, u, true);
    GatherStrided(a, acc, b, e, 1);
    r += Finish(a, u, static_cast<T>(acc.Value()));
  }
  return r;
}

template <class T>
__device__ __forceinline__ double GatherWavePart(const PrArgs<T>& a, const int32_t* __restrict__ verts, int32_t nv,
                                                int64_t blk, int64_t nblk) {
  using A = typename Acc<T>::Type;
  const int lane = LaneId();
  const int64_t wave = (blk * kBS + threadIdx.x) / kWave;
  const int64_t nwaves = nblk * (kBS / kWave);
  double r = 0.0;
  for (int64_t i = wave; i < nv; i += nwaves) {
    const int32_t u = verts[i];
    if (u < a.lo || u >= a.hi) continue;  // uniform per wavefront
    const int64_t b = ListBegin(a, u), e = a.in_off[u + 1];
    A acc = StartSum(a, u, lane == 0);
    GatherStrided(a, acc, b + lane, e, kWave);
    acc = WaveMerge(acc);
    if (lane == 0) r += Finish(a, u, static_cast<T>(acc.Value()));
  }
  return r;
}

template <class T>
__device__ __forceinline__ double GatherBlockPart(const PrArgs<T>& a, const int32_t* __restrict__ verts, int32_t nv,
                                                 int64_t blk, int64_t nblk,
                                                 typename Acc<T>::Type* s_acc) {
  using A = typename Acc<T>::Type;
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  double r = 0.0;
  for (int64_t i = blk; i < nv; i += nblk) {
    const int32_t u = verts[i];
    if (u < a.lo || u >= a.hi) continue;  // uniform per block
    const int64_t b = ListBegin(a, u), e = a.in_off[u + 1];
    A acc = StartSum(a, u, threadIdx.x == 0);
    GatherStrided(a, acc, b + threadIdx.x, e, kBS);
    acc = WaveMerge(acc);
    if (lane == 0) s_acc[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      A t = s_acc[0];
      for (int w = 1; w < kBS / kWave; w++) t.Merge(s_acc[w]);
      r += Finish(a, u, static_cast<T>(t.Value()));
    }
    __syncthreads();
  }
  return r;
}

template <class T>
__global__ __launch_bounds__(kBS) void PrGather(PrArgs<T> a, const int32_t* __restrict__ small, int32_t n_small,
                                                const int32_t* __restrict__ medium, int32_t n_medium,
                                                const int32_t* __restrict__ large, int32_t n_large,
                                                int32_t nb_large, int32_t nb_medium) {
  __shared__ double s_red[kBS / kWave];
  __shared__ typename Acc<T>::Type s_acc[kBS / kWave];
  const int64_t b = blockIdx.x;
  double r;
  if (b < nb_large)
    r = GatherBlockPart(a, large, n_large, b, nb_large, s_acc);
  else if (b < nb_large + nb_medium)
    r = GatherWavePart(a, medium, n_medium, b - nb_large, nb_medium);
  else
    r = GatherThreadPart(a, small, n_small, b - nb_large - nb_medium, int64_t(gridDim.x) - nb_large - nb_medium);
  AddResidual(a.residual, r, s_red);
}

template <class T>
__global__ __launch_bounds__(kBS) void PrInit(const int64_t* __restrict__ out_off, T* __restrict__ score,
                                              T* __restrict__ contrib, typename Acc<T>::Stored* __restrict__ acc,
                                              int64_t n, T init) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    score[i] = init;
    contrib[i] = init / static_cast<T>(out_off[i + 1] - out_off[i]);
    if (acc) acc[i] = typename Acc<T>::Stored{};
  }
}

// ---------------------------------------------------------------------------
// Source segments.

// One segment launch: every pair of segment s adds the sum of its sources'
// contributions to its destination's accumulator.  A destination has at
// most one pair per segment, so the read-modify-write of acc needs no
// atomics; launches of different segments are ordered by the stream.
template <class T>
struct SegArgs {
  const T* contrib;
  const int32_t* src;                // sub-lists, segment-major
  const int32_t* dst;                // destination of each pair
  const int64_t* off;                // pair p's sources: src[off[p], off[p + 1])
  typename Acc<T>::Stored* acc;
  int64_t p0;                        // the segment's first pair
  int32_t n_large, n_medium, n_small;  // its pairs per bin, stored in this order
  int32_t nb_large, nb_medium;       // blocks serving the block and wavefront bins
};

template <class T>
__device__ __forceinline__ void AddToAcc(typename Acc<T>::Stored* acc, int32_t v, const typename Acc<T>::Type& part) {
  auto t = Unpack(acc[v]);
  t.Merge(part);
  acc[v] = Pack(t);
}

// Pair bins of the segment layout, by sub-list length: a block per pair
// from kPairBlock entries, a wavefront from kPairWave, else a thread.
constexpr int kPairWave = 32;
constexpr int kPairBlock = 1024;

template <class T>
__global__ __launch_bounds__(kBS) void PrSegment(SegArgs<T> a) {
  using A = typename Acc<T>::Type;
  __shared__ A s_acc[kBS / kWave];
  const int64_t b = blockIdx.x;
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  if (b < a.nb_large) {  // one block per pair
    for (int64_t i = b; i < a.n_large; i += a.nb_large) {
      const int64_t p = a.p0 + i;
      A acc = A::Zero();
      GatherIds(a.contrib, a.src, acc, a.off[p] + threadIdx.x, a.off[p + 1], kBS);
      acc = WaveMerge(acc);
      if (lane == 0) s_acc[wid] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        A t = s_acc[0];
        for (int w = 1; w < kBS / kWave; w++) t.Merge(s_acc[w]);
        AddToAcc<T>(a.acc, a.dst[p], t);
      }
      __syncthreads();
    }
  } else if (b < a.nb_large + a.nb_medium) {  // one wavefront per pair
    const int64_t nw = int64_t(a.nb_medium) * (kBS / kWave);
    for (int64_t i = (b - a.nb_large) * (kBS / kWave) + wid; i < a.n_medium; i += nw) {
      const int64_t p = a.p0 + a.n_large + i;
      A acc = A::Zero();
      GatherIds(a.contrib, a.src, acc, a.off[p] + lane, a.off[p + 1], kWave);
      acc = WaveMerge(acc);
      if (lane == 0) AddToAcc<T>(a.acc, a.dst[p], acc);
    }
  } else {  // one thread per pair
    const int64_t nb = int64_t(gridDim.x) - a.nb_large - a.nb_medium;
    for (int64_t i = (b - a.nb_large - a.nb_medium) * kBS + threadIdx.x; i < a.n_small; i += nb * kBS) {
      const int64_t p = a.p0 + a.n_large + a.n_medium + i;
      A acc = A::Zero();
      GatherIds(a.contrib, a.src, acc, a.off[p], a.off[p + 1], 1);
      AddToAcc<T>(a.acc, a.dst[p], acc);
    }
  }
}

// LDS segments: a block of kLdsBS threads per CU stages the segment's
// contributions (kLdsBytes) in LDS and gathers from there.
constexpr int kLdsBS = 1024;
constexpr int kLdsBytes = 128 * 1024;

template <class T>
__device__ __forceinline__ void GatherLds(const T* __restrict__ slice, int64_t base, const int32_t* __restrict__ ids,
                                          typename Acc<T>::Type& acc, int64_t k0, int64_t e, int step) {
  int64_t k = k0;
  for (; k + 3 * step < e; k += 4 * step) {
    const int32_t v0 = ids[k], v1 = ids[k + step], v2 = ids[k + 2 * step], v3 = ids[k + 3 * step];
    const T x0 = slice[v0 - base], x1 = slice[v1 - base], x2 = slice[v2 - base], x3 = slice[v3 - base];
    acc.Add(x0);
    acc.Add(x1);
    acc.Add(x2);
    acc.Add(x3);
  }
  for (; k < e; k += step) acc.Add(slice[ids[k] - base]);
}

// One LDS segment launch, one block per CU: the segment's contributions
// [base, base + count) are staged in LDS, then the block takes its share of
// the segment's block, wavefront and short pairs in turn.
template <class T>
__global__ __launch_bounds__(kLdsBS) void PrSegmentLds(SegArgs<T> a, int64_t base, int32_t count) {
  using A = typename Acc<T>::Type;
  __shared__ T slice[kLdsBytes / sizeof(T)];
  __shared__ A s_acc[kLdsBS / kWave];
  for (int i = threadIdx.x; i < count; i += kLdsBS) slice[i] = a.contrib[base + i];
  __syncthreads();
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  for (int64_t i = blockIdx.x; i < a.n_large; i += gridDim.x) {
    const int64_t p = a.p0 + i;
    A acc = A::Zero();
    GatherLds<T>(slice, base, a.src, acc, a.off[p] + threadIdx.x, a.off[p + 1], kLdsBS);
    acc = WaveMerge(acc);
    if (lane == 0) s_acc[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      A t = s_acc[0];
      for (int w = 1; w < kLdsBS / kWave; w++) t.Merge(s_acc[w]);
      AddToAcc<T>(a.acc, a.dst[p], t);
    }
    __syncthreads();
  }
  const int64_t nw = int64_t(gridDim.x) * (kLdsBS / kWave);
  for (int64_t i = int64_t(blockIdx.x) * (kLdsBS / kWave) + wid; i < a.n_medium; i += nw) {
    const int64_t p = a.p0 + a.n_large + i;
    A acc = A::Zero();
    GatherLds<T>(slice, base, a.src, acc, a.off[p] + lane, a.off[p + 1], kWave);
    acc = WaveMerge(acc);
    if (lane == 0) AddToAcc<T>(a.acc, a.dst[p], acc);
  }
  // Short pairs, two per thread at a time (their loads interleaved).
  const int64_t q0 = a.p0 + a.n_large + a.n_medium, stride = int64_t(gridDim.x) * kLdsBS;
  for (int64_t i = int64_t(blockIdx.x) * kLdsBS + threadIdx.x; i < a.n_small; i += 2 * stride) {
    const int64_t p = q0 + i;
    const bool two = i + stride < a.n_small;
    const int64_t pb = two ? p + stride : p;
    const int64_t b0 = a.off[p], e0 = a.off[p + 1], b1 = a.off[pb], e1 = two ? a.off[pb + 1] : b1;
    const int32_t d0 = a.dst[p], d1 = a.dst[pb];
    A acc0 = A::Zero(), acc1 = A::Zero();
    const int64_t len = e0 - b0 > e1 - b1 ? e0 - b0 : e1 - b1;
    for (int64_t k = 0; k < len; k++) {
      const bool h0 = b0 + k < e0, h1 = b1 + k < e1;
      const int32_t v0 = h0 ? a.src[b0 + k] : 0, v1 = h1 ? a.src[b1 + k] : 0;
      if (h0) acc0.Add(slice[v0 - base]);
      if (h1) acc1.Add(slice[v1 - base]);
    }
    AddToAcc<T>(a.acc, d0, acc0);
    if (two) AddToAcc<T>(a.acc, d1, acc1);
  }
}

// Building the layout (once per graph and precision, outside timing).
constexpr int kLdsSegMax = 128;            // candidate LDS segments examined
constexpr int kL2SegMax = 64;              // candidate L2 segments examined
constexpr int kSegMax = kLdsSegMax + kL2SegMax;
constexpr int kSegMinEntriesPerPair = 3;   // density a default-width segment needs (gg_tuning can change it)
constexpr int kCopyTile = 4096;            // output entries per block of the sub-list copy

// Segment geometry: n_lds segments of 2^sbl sources, then segments of 2^sb2
// sources up to ns in all.
struct SegGeom {
  int sbl, sb2, n_lds, ns;
  __device__ __forceinline__ int64_t LdsEnd() const { return int64_t(n_lds) << sbl; }
  __device__ __forceinline__ int64_t Hot() const { return LdsEnd() + (int64_t(ns - n_lds) << sb2); }
  __device__ __forceinline__ int Of(int64_t x) const {
    return x < LdsEnd() ? static_cast<int>(x >> sbl) : n_lds + static_cast<int>((x - LdsEnd()) >> sb2);
  }
  __device__ __forceinline__ int64_t End(int s) const {
    return s < n_lds ? int64_t(s + 1) << sbl : LdsEnd() + (int64_t(s + 1 - n_lds) << sb2);
  }
};

// End of the run of sources below x that starts at b (a sorted list ending
// at e): a short linear probe first, since most runs are short.
__device__ __forceinline__ int64_t RunEnd(const int32_t* __restrict__ dst, int64_t b, int64_t e, int64_t x) {
  const int64_t probe_end = b + 8 < e ? b + 8 : e;
  for (; b < probe_end; b++)
    if (dst[b] >= x) return b;
  while (b < e) {
    const int64_t m = b + ((e - b) >> 1);
    if (dst[m] < x)
      b = m + 1;
    else
      e = m;
  }
  return b;
}

// Flags a list that is not sorted, or too long for a 32-bit tail offset
// (the layout then stays off).  One wavefront per vertex.
__global__ __launch_bounds__(kBS) void PlCheck(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                               int64_t n, int32_t* bad) {
  const int lane = LaneId();
  const int64_t nw = int64_t(gridDim.x) * (kBS / kWave);
  for (int64_t v = (int64_t(blockIdx.x) * kBS + threadIdx.x) / kWave; v < n; v += nw) {
    const int64_t b = in_off[v], e = in_off[v + 1];
    bool wrong = e - b > INT32_MAX;
    for (int64_t k = b + lane; !wrong && k + 1 < e; k += kWave) wrong = in_dst[k] > in_dst[k + 1];
    if (__ballot(wrong) != 0 && lane == 0) atomicOr(bad, 1);
  }
}

struct PlTotals {
  unsigned long long pairs[kSegMax][3];  // per segment and bin (block, wavefront, thread)
  unsigned long long entries[kSegMax];
  int32_t bad;
};

// Pass 1: for each tile of kBS vertices, the pairs it holds per segment and
// bin (stored (segment, bin)-major, tile-minor, so one scan turns them into
// output positions), and the totals per segment.
__global__ __launch_bounds__(kBS) void PlCount(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                               int64_t n, SegGeom G, int32_t small_max, int32_t medium_max,
                                               int64_t tiles, uint32_t* __restrict__ counts, PlTotals* tot) {
  __shared__ uint32_t s_cnt[kSegMax * 3];
  __shared__ unsigned long long s_ent[kSegMax];
  for (int i = threadIdx.x; i < kSegMax * 3; i += kBS) s_cnt[i] = 0;
  for (int i = threadIdx.x; i < kSegMax; i += kBS) s_ent[i] = 0;
  __syncthreads();
  const int64_t tile = blockIdx.x;
  const int64_t v = tile * kBS + threadIdx.x;
  if (v < n) {
    const int64_t hot = G.Hot();
    const int64_t end = in_off[v + 1];
    int64_t pos = in_off[v];
    while (pos < end) {
      const int64_t x = in_dst[pos];
      if (x >= hot) break;
      const int s = G.Of(x);
      const int64_t nxt = RunEnd(in_dst, pos + 1, end, G.End(s));
      const int64_t len = nxt - pos;
      const int bin = len >= medium_max ? 0 : (len >= small_max ? 1 : 2);
      atomicAdd(&s_cnt[s * 3 + bin], 1u);
      atomicAdd(&s_ent[s], static_cast<unsigned long long>(len));
      pos = nxt;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < G.ns * 3; i += kBS) {
    counts[int64_t(i) * tiles + tile] = s_cnt[i];
    if (s_cnt[i]) atomicAdd(&tot->pairs[i / 3][i % 3], static_cast<unsigned long long>(s_cnt[i]));
  }
  for (int s = threadIdx.x; s < G.ns; s += kBS)
    if (s_ent[s]) atomicAdd(&tot->entries[s], s_ent[s]);
}

// Pass 2: the same walk over the first ns segments, writing each pair's
// destination, length and list position at its place (segment, then bin,
// then tile, then vertex order), and each vertex's tail offset.
__global__ __launch_bounds__(kBS) void PlEmit(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                              int64_t n, SegGeom G, int32_t small_max, int32_t medium_max,
                                              int64_t tiles, const uint32_t* __restrict__ base,
                                              int32_t* __restrict__ pair_dst, int64_t* __restrict__ pair_len,
                                              int64_t* __restrict__ pair_pos, int32_t* __restrict__ tail) {
  __shared__ int sh[kBS / kWave];
  const int64_t tile = blockIdx.x;
  const int64_t v = tile * kBS + threadIdx.x;
  const int64_t hot = G.Hot();
  int64_t beg = 0, pos = 0, end = 0;
  int cur = G.ns;  // segment of the next run, ns once none is left
  if (v < n) {
    beg = pos = in_off[v];
    end = in_off[v + 1];
    if (pos < end && in_dst[pos] < hot) cur = G.Of(in_dst[pos]);
  }
  for (int s = 0; s < G.ns; s++) {
    const bool has = cur == s;
    int64_t nxt = pos;
    int bin = 0;
    if (has) {
      nxt = RunEnd(in_dst, pos + 1, end, G.End(s));
      const int64_t len = nxt - pos;
      bin = len >= medium_max ? 0 : (len >= small_max ? 1 : 2);
    }
    int total;  // per-bin ranks in 10-bit fields (at most kBS per field)
    const int excl = BlockExclusiveScan<kBS>(has ? 1 << (10 * bin) : 0, &total, sh);
    if (has) {
      const int64_t p = int64_t(base[int64_t(s * 3 + bin) * tiles + tile]) + ((excl >> (10 * bin)) & 1023);
      pair_dst[p] = static_cast<int32_t>(v);
      pair_len[p] = nxt - pos;
      pair_pos[p] = pos;
      pos = nxt;
      cur = pos < end && in_dst[pos] < hot ? G.Of(in_dst[pos]) : G.ns;
    }
  }
  if (v < n) tail[v] = static_cast<int32_t>(pos - beg);
}

// First index in [lo, hi) with a[index] > x.
__device__ __forceinline__ int64_t UpperBound(const int64_t* __restrict__ a, int64_t lo, int64_t hi, int64_t x) {
  while (lo < hi) {
    const int64_t m = lo + ((hi - lo) >> 1);
    if (a[m] <= x)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

// Pass 3: copies every pair's sub-list to its place in the segment-major
// array, kCopyTile outputs per block (owner pair by binary search).
__global__ __launch_bounds__(kBS) void PlCopy(const int64_t* __restrict__ pair_off, const int64_t* __restrict__ pair_pos,
                                              int64_t pairs, const int32_t* __restrict__ in_dst,
                                              int32_t* __restrict__ src, int64_t total) {
  __shared__ int64_t s_lo, s_hi;
  for (int64_t b0 = int64_t(blockIdx.x) * kCopyTile; b0 < total; b0 += int64_t(gridDim.x) * kCopyTile) {
    const int64_t b1 = b0 + kCopyTile < total ? b0 + kCopyTile : total;
    if (threadIdx.x == 0) {
      s_lo = UpperBound(pair_off, 0, pairs + 1, b0) - 1;
      s_hi = UpperBound(pair_off, 0, pairs + 1, b1 - 1);
    }
    __syncthreads();
    const int64_t lo = s_lo, hi = s_hi;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kBS) {
      const int64_t p = UpperBound(pair_off, lo, hi, i) - 1;
      src[i] = in_dst[pair_pos[p] + (i - pair_off[p])];
    }
    __syncthreads();
  }
}

__global__ void PlSetLast(int64_t* pair_off, int64_t pairs, int64_t total) { pair_off[pairs] = total; }

__global__ __launch_bounds__(kBS) void PrResetResidual(double* residual) {
  if (threadIdx.x < kStripes) residual[threadIdx.x] = 0.0;
}

// Posts the summed residual to the host mailbox (as raw bits) and clears it.
__global__ __launch_bounds__(kBS) void PrPublish(double* residual, unsigned long long* mb,
                                                 unsigned long long seq) {
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kStripes; i++) {
      t += residual[i];
      residual[i] = 0.0;
    }
    unsigned long long bits = __double_as_longlong(t);
    MailPost(mb, &bits, 1, seq);
  }
}

__global__ __launch_bounds__(kBS) void ToFloat(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    out[i] = static_cast<float>(in[i]);
}

// Bins vertices by in-degree (once per graph); also sums each bin's edges.
__global__ __launch_bounds__(kBS) void PrBin(const int64_t* __restrict__ in_off, int64_t n,
                                             int32_t* __restrict__ small, int32_t* __restrict__ medium,
                                             int32_t* __restrict__ large, int32_t* tails,
                                             unsigned long long* edges, int64_t small_max, int64_t medium_max) {
  __shared__ unsigned long long s_red[kBS / kWave];
  unsigned long long e[3] = {0, 0, 0};
  for (int64_t base = int64_t(blockIdx.x) * kBS; base < n; base += int64_t(gridDim.x) * kBS) {
    const int64_t u = base + threadIdx.x;
    int64_t d = -1;
    if (u < n) d = in_off[u + 1] - in_off[u];
    const int32_t v = static_cast<int32_t>(u);
    const int bin = d < 0 ? -1 : (d < small_max ? 0 : (d < medium_max ? 1 : 2));
    if (bin >= 0) e[bin] += static_cast<unsigned long long>(d);
    WaveAppend(bin == 0, v, small, &tails[0]);
    WaveAppend(bin == 1, v, medium, &tails[1]);
    WaveAppend(bin == 2, v, large, &tails[2]);
  }
  for (int k = 0; k < 3; k++) {
    const unsigned long long t = BlockSum<kBS>(e[k], s_red);
    if (threadIdx.x == 0 && t) atomicAdd(&edges[k], t);
  }
}

struct PrBins {
  int32_t *small, *medium, *large;
  int64_t n_small, n_medium, n_large;
  int64_t e_small, e_medium, e_large;
};

PrBins GetBins(gg_graph* g) {
  PrBins b;
  b.small = g->Scratch<int32_t>("pr.bin_small", g->n + 1);
  b.medium = g->Scratch<int32_t>("pr.bin_medium", g->n + 1);
  b.large = g->Scratch<int32_t>("pr.bin_large", g->n + 1);
  if (g->pr_bins[0] < 0) {
    struct Tally {
      int32_t tails[4];
      unsigned long long edges[3];
    };
    Tally* d = g->Scratch<Tally>("pr.bin_tally", 1);
    GG_HIP(hipMemsetAsync(d, 0, sizeof(Tally), g->stream));
    PrBin<<<GridFor(g->n, kBS, g->cu_count * 8), kBS, 0, g->stream>>>(
        g->in_off, g->n, b.small, b.medium, b.large, d->tails, d->edges, g->tune.pr_small_max,
        g->tune.pr_medium_max);
    Tally h;
    GG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    for (int i = 0; i < 3; i++) {
      g->pr_bins[i] = h.tails[i];
      g->pr_bin_edges[i] = static_cast<int64_t>(h.edges[i]);
    }
  }
  b.n_small = g->pr_bins[0];
  b.n_medium = g->pr_bins[1];
  b.n_large = g->pr_bins[2];
  b.e_small = g->pr_bin_edges[0];
  b.e_medium = g->pr_bin_edges[1];
  b.e_large = g->pr_bin_edges[2];
  return b;
}

// The source-segment layout of precision T, built on first use (and again
// when a width or the density rule changes); null when the gather runs
// without segments.  Segments of two kinds, in source order: LDS segments
// (only when gg_tuning.pr_lds_log2 asks for them), whose launches stage up
// to kLdsBytes of contributions in LDS, then L2 segments of 2 MB (2^19 float
// or 2^18 double sources by default), whose contributions stay in every
// XCD's L2.  By default (graphs of 2^20 vertices or more) the leading L2
// segments with at least kSegMinEntriesPerPair entries per pair (or the
// tuned density) are kept; explicit widths take every candidate segment.
// Lists must be sorted (they are for builder output and degree-ordered
// handles).
template <class T>
const gg_graph::PrSegments* Segments(gg_graph* g) {
  const int prec = sizeof(T) == 4 ? 0 : 1;
  const int32_t knob = g->tune.pr_segment_log2[prec], lknob = g->tune.pr_lds_log2[prec];
  if (knob < 0) return nullptr;
  gg_graph::PrSegments& L = g->pr_seg[prec];
  const int sb2 = knob > 0 ? knob : (sizeof(T) == 4 ? 19 : 18);
  const int sbl = lknob > 0 ? lknob : -1;
  const int32_t density = g->tune.pr_seg_min_density > 0 ? g->tune.pr_seg_min_density : kSegMinEntriesPerPair;
  if (L.built && L.sb == sb2 && L.sbl == sbl && L.min_density == density) return L.ns > 0 ? &L : nullptr;
  const std::string tag = sizeof(T) == 4 ? "pr.seg32." : "pr.seg64.";
  g->ws.Drop(tag);
  WallTimer wt;
  L = gg_graph::PrSegments{};
  L.built = true;
  L.sb = sb2;
  L.sbl = sbl;
  L.small_max = kPairWave;
  L.medium_max = kPairBlock;
  L.min_density = density;
  const int64_t n = g->n;
  if (knob == 0 && sbl < 0 && n < (int64_t(1) << 20)) return nullptr;
  const bool prof = g->profile;
  g->profile = false;  // the build is not part of any measured call
  struct Restore {
    gg_graph* g;
    bool prof;
    ~Restore() { g->profile = prof; }
  } restore{g, prof};
  hipStream_t st = g->stream;
  const int64_t tiles = (n + kBS - 1) / kBS;
  PlTotals* dtot = g->Scratch<PlTotals>("pr.segb.tot", 1);
  uint32_t* counts = g->Scratch<uint32_t>("pr.segb.counts", static_cast<size_t>(kSegMax) * 3 * tiles);
  GG_HIP(hipMemsetAsync(dtot, 0, sizeof(PlTotals), st));
  PlCheck<<<GridFor(n, kBS / kWave, g->cu_count * 16), kBS, 0, st>>>(g->in_off, g->in_dst, n, &dtot->bad);
  PlTotals tot;
  auto count = [&](const SegGeom& G) {
    GG_HIP(hipMemsetAsync(dtot->pairs, 0, sizeof(tot.pairs) + sizeof(tot.entries), st));
    PlCount<<<static_cast<unsigned>(tiles), kBS, 0, st>>>(g->in_off, g->in_dst, n, G, L.small_max, L.medium_max,
                                                          tiles, counts, dtot);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(&tot, dtot, sizeof(tot), hipMemcpyDeviceToHost, st));
    GG_HIP(hipStreamSynchronize(st));
  };
  // Leading segments of [s0, s1) dense enough to keep (all of them with an
  // explicit width).
  auto keep = [&](int s0, int s1, bool all) {
    int s = s0;
    for (; s < s1; s++) {
      const unsigned long long p = tot.pairs[s][0] + tot.pairs[s][1] + tot.pairs[s][2];
      if (p == 0 || (!all && tot.entries[s] < static_cast<unsigned long long>(density) * p)) break;
    }
    return s;
  };
  SegGeom G{sbl < 0 ? 0 : sbl, sb2, 0, 0};
  if (sbl >= 0) {
    const int cand = static_cast<int>(std::min<int64_t>(kLdsSegMax, ((n - 1) >> sbl) + 1));
    G.n_lds = G.ns = cand;
    count(G);
    if (tot.bad) {
      g->ws.Drop("pr.segb.");
      return nullptr;
    }
    G.n_lds = G.ns = keep(0, cand, lknob > 0);
  }
  const int64_t rest = n - (int64_t(G.n_lds) << G.sbl);
  const int cand2 = rest > 0 ? static_cast<int>(std::min<int64_t>(kL2SegMax, ((rest - 1) >> sb2) + 1)) : 0;
  G.ns = G.n_lds + cand2;
  if (G.ns > 0) count(G);
  if (G.ns == 0 || tot.bad) {
    g->ws.Drop("pr.segb.");
    return nullptr;
  }
  G.ns = keep(G.n_lds, G.n_lds + cand2, knob > 0);
  int64_t pairs = 0, hot_entries = 0;
  for (int s = 0; s < G.ns; s++) {
    pairs += static_cast<int64_t>(tot.pairs[s][0] + tot.pairs[s][1] + tot.pairs[s][2]);
    hot_entries += static_cast<int64_t>(tot.entries[s]);
  }
  if (G.ns == 0 || pairs == 0 || pairs >= (int64_t(1) << 31)) {  // pair counts and positions are 32-bit
    g->ws.Drop("pr.segb.");
    return nullptr;
  }
  L.ns = G.ns;
  L.n_lds = G.n_lds;
  L.hot = std::min<int64_t>((int64_t(G.n_lds) << G.sbl) + (int64_t(G.ns - G.n_lds) << sb2), n);
  L.pairs = pairs;
  L.hot_entries = hot_entries;
  L.first_pair.assign(G.ns + 1, 0);
  L.large.resize(G.ns);
  L.medium.resize(G.ns);
  L.entries.resize(G.ns);
  for (int s = 0; s < G.ns; s++) {
    L.large[s] = static_cast<int64_t>(tot.pairs[s][0]);
    L.medium[s] = static_cast<int64_t>(tot.pairs[s][1]);
    L.entries[s] = static_cast<int64_t>(tot.entries[s]);
    L.first_pair[s + 1] = L.first_pair[s] + L.large[s] + L.medium[s] + static_cast<int64_t>(tot.pairs[s][2]);
  }
  // Output positions of each (segment, bin, tile) run of pairs (the counts
  // of the kept segments lead the array).
  DeviceScan<uint32_t>(g, "pr.segb.scan", counts, int64_t(G.ns) * 3 * tiles, nullptr);
  int32_t* pair_dst = g->Scratch<int32_t>((tag + "dst").c_str(), pairs);
  int64_t* pair_off = g->Scratch<int64_t>((tag + "off").c_str(), pairs + 1);
  int32_t* tail = g->Scratch<int32_t>((tag + "tail").c_str(), n);
  int32_t* src = g->Scratch<int32_t>((tag + "src").c_str(), hot_entries);
  int64_t* pair_pos = g->Scratch<int64_t>("pr.segb.pos", pairs);
  PlEmit<<<static_cast<unsigned>(tiles), kBS, 0, st>>>(g->in_off, g->in_dst, n, G, L.small_max, L.medium_max, tiles,
                                                       counts, pair_dst, pair_off, pair_pos, tail);
  DeviceScan<int64_t>(g, "pr.segb.scan64", pair_off, pairs, nullptr);
  PlSetLast<<<1, 1, 0, st>>>(pair_off, pairs, hot_entries);
  PlCopy<<<GridFor(hot_entries, kCopyTile, g->cu_count * 16), kBS, 0, st>>>(pair_off, pair_pos, pairs, g->in_dst, src,
                                                                           hot_entries);
  GG_HIP(hipGetLastError());
  GG_HIP(hipStreamSynchronize(st));
  g->ws.Drop("pr.segb.");
  L.build_ms = wt.ms();
  return &L;
}

// ---------------------------------------------------------------------------
// Propagation blocks (Jacobi modes; the default on graphs of 2^20 to 2^27
// vertices).
//
// After Beamer, Asanovic and Patterson, "Reducing PageRank Communication via
// Propagation Blocking" (IPDPS 2017).  A pull gather reads contrib[v] at
// random for every in-entry, and a random read costs one cache-line
// transaction whether it wants 4 or 8 bytes.  Here an iteration instead
// writes every in-entry's contribution into a value array whose order lets
// both phases stream.  The in-entries are grouped into blocks (t, b):
// source in tile t, destination in bin b, both 2^lg consecutive ids.  The
// blocks are stored tile-major, and each block's entries are kept together
// by destination and padded to a multiple of 8 (an octet) with (source
// offset 0, destination 0xFFFF).  Ids are stored as 16-bit offsets within
// their tile or bin.
//   write  a block of threads per chunk of one tile's octets: the tile's
//          contributions are staged in LDS, then the chunk's source offsets
//          (2 B per entry) are streamed and each entry's contribution
//          (4 or 8 B) is written in the same order;
//   sum    a block per chunk of one bin's octets, read through an octet
//          index kept bin-major (4 B per octet): each value is added to its
//          destination's accumulator in LDS, then the bin's vertices are
//          finished (or, for a bin split into several chunks, the chunk's
//          partial sums are stored);
//   merge  for the bins of several chunks: their partial sums are merged
//          in chunk order and the vertices finished.
// Bytes per in-entry: 2 + 4 + 0.5 + 2 + 4 = 12.5 in FP32 and 20.5 in FP64,
// so the contribution traffic scales with the element size.
// Accumulators (8 bytes per destination in LDS, added to with atomics by
// the block's lanes):
//   FP32  an exact sum in 64-bit fixed point, units of 2^-63.  Every
//         contribution is score / out-degree <= 1, and a vertex's sum is at
//         most the sum of all scores, which PageRank without dangling
//         redistribution keeps at or below 1, so sums fit below 2.  A value
//         is rounded once, to the nearest unit (an absolute error of at
//         most 2^-64 per entry), and the rest is integer arithmetic:
//         exact, the same in any order, and one 64-bit atomic per run of
//         equal destinations.  The finished sum is rounded to float once.
//         (A compensated float pair, the other FP32 paths' accumulator,
//         would take two atomics per run here, one of them returning the
//         old value for its TwoSum, and measured 2.2x slower at Kronecker
//         scale 26.)
//   FP64  a double, added in arrival order.
constexpr int kPbMaxLog = 14;      // 2^14 accumulators of 8 bytes fill 128 KB of LDS
constexpr int kPbMaxTiles = 8192;  // build tables hold k x k octet counts
constexpr int kPbWriteBS = 1024;
constexpr int kPbSumBS = 1024;
constexpr int kPbBuildBS = 1024;

// The per-destination accumulator word of each precision.
template <class T>
struct PbAcc;
template <>
struct PbAcc<float> {
  using Word = unsigned long long;  // 64-bit fixed point, units of 2^-63
};
template <>
struct PbAcc<double> {
  using Word = double;
};


template <class T>
struct PbArgs {
  const T* contrib;                 // this iteration's contributions
  T* val;                           // one value per (padded) entry, tile-major
  const uint4* src16;               // per octet: 8 source offsets within the tile
  const uint4* dst16;               // per octet: 8 destination offsets within the bin
  const uint32_t* oct;              // each bin's octets, bin-major
  typename PbAcc<T>::Word* part;    // partial sums, 2^lg per slot
  const int64_t* out_off;
  T* score;
  T* contrib_out;
  T base, damp;
  double* residual;
  int64_t n;
  int lg;
};

// Entries per lane in the write phase: 4 floats or 2 doubles, so that each
// lane stores 16 contiguous bytes and a wavefront's store is one contiguous
// kilobyte.
template <class T>
struct PbWriteShape;
template <>
struct PbWriteShape<float> {
  static constexpr int kEntries = 4, kItems = 1;  // entries per item, items in flight per thread
  using Ids = uint2;
  using Vals = float4;
  __device__ static void Unpack(const uint2 w, uint32_t (&o)[4]) {
    o[0] = w.x & 0xFFFFu, o[1] = w.x >> 16, o[2] = w.y & 0xFFFFu, o[3] = w.y >> 16;
  }
  __device__ static float4 Pack(const float (&v)[4]) { return make_float4(v[0], v[1], v[2], v[3]); }
};
template <>
struct PbWriteShape<double> {
  static constexpr int kEntries = 2, kItems = 2;
  using Ids = uint32_t;
  using Vals = double2;
  __device__ static void Unpack(const uint32_t w, uint32_t (&o)[2]) { o[0] = w & 0xFFFFu, o[1] = w >> 16; }
  __device__ static double2 Pack(const double (&v)[2]) { return make_double2(v[0], v[1]); }
};

__device__ __forceinline__ void LoadOctet(const float* val, int64_t q, float (&v)[8]) {
  const float4* p = reinterpret_cast<const float4*>(val) + 2 * q;
  const float4 a = p[0], b = p[1];
  v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
}
__device__ __forceinline__ void LoadOctet(const double* val, int64_t q, double (&v)[8]) {
  const double2* p = reinterpret_cast<const double2*>(val) + 4 * q;
  const double2 a = p[0], b = p[1], c = p[2], d = p[3];
  v[0] = a.x, v[1] = a.y, v[2] = b.x, v[3] = b.y, v[4] = c.x, v[5] = c.y, v[6] = d.x, v[7] = d.y;
}
__device__ __forceinline__ void Unpack16(const uint4 w, uint32_t (&o)[8]) {
  o[0] = w.x & 0xFFFFu, o[1] = w.x >> 16, o[2] = w.y & 0xFFFFu, o[3] = w.y >> 16;
  o[4] = w.z & 0xFFFFu, o[5] = w.z >> 16, o[6] = w.w & 0xFFFFu, o[7] = w.w >> 16;
}

// A contribution (0 <= x <= 1) in fixed-point units of 2^-63, rounded to
// the nearest unit; and a fixed-point sum back to float (one rounding).
__device__ __forceinline__ unsigned long long ToFixed(float x) { return __float2ull_rn(__fmul_rn(x, 0x1p63f)); }
__device__ __forceinline__ float FromFixed(unsigned long long w) { return __fmul_rn(__ull2float_rn(w), 0x1p-63f); }

// Adds one octet's entries to the LDS accumulators: each run of equal
// destinations is summed in registers and added at its last entry.
// d: the 8 destination offsets (kPbPad: padding, only after real entries).
__device__ __forceinline__ void PbAddOctet(unsigned long long* acc, const uint32_t (&d)[8], const float (&v)[8],
                                           uint32_t m) {
  unsigned long long run = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    if (d[i] == kPbPad) break;
    run = (i > 0 && d[i] != d[i - 1] ? 0ull : run) + ToFixed(v[i]);
    if (i == 7 || d[i + 1] != d[i]) atomicAdd(&acc[d[i] & m], run);
  }
}
__device__ __forceinline__ void PbAddOctet(double* acc, const uint32_t (&d)[8], const double (&v)[8], uint32_t m) {
  double run = 0.0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    if (d[i] == kPbPad) break;
    run = i > 0 && d[i] != d[i - 1] ? v[i] : __dadd_rn(run, v[i]);
    if (i == 7 || d[i + 1] != d[i]) atomicAdd(&acc[d[i] & m], run);
  }
}
__device__ __forceinline__ float PbValue(unsigned long long w) { return FromFixed(w); }
__device__ __forceinline__ double PbValue(double w) { return w; }

template <class T>
__device__ __forceinline__ double PbFinish(const PbArgs<T>& a, int64_t u, T sum) {
  const T old = a.score[u];
  const T now = Combine(a.base, a.damp, sum);
  a.score[u] = now;
  const int64_t deg = a.out_off[u + 1] - a.out_off[u];
  a.contrib_out[u] = now / static_cast<T>(deg);
  return fabs(static_cast<double>(now) - static_cast<double>(old));
}

template <class T>
__global__ __launch_bounds__(kPbWriteBS) void PbWrite(PbArgs<T> a, const PbChunk* __restrict__ chunks) {
  using S = PbWriteShape<T>;
  constexpr int E = S::kEntries, U = S::kItems;
  __shared__ T slice[1 << kPbMaxLog];
  const PbChunk c = chunks[blockIdx.x];
  const int64_t base = int64_t(c.part) << a.lg;
  const int64_t cnt = min(int64_t(1) << a.lg, a.n - base);
  for (int64_t i = threadIdx.x; i < cnt; i += kPbWriteBS) slice[i] = a.contrib[base + i];
  __syncthreads();
  const uint32_t m = (1u << a.lg) - 1;  // keeps every slice index below 2^lg
  const typename S::Ids* ids = reinterpret_cast<const typename S::Ids*>(a.src16);
  typename S::Vals* out = reinterpret_cast<typename S::Vals*>(a.val);
  const int64_t i1 = int64_t(c.q1) * (8 / E);  // items of E entries
  for (int64_t i = int64_t(c.q0) * (8 / E) + threadIdx.x; i < i1; i += int64_t(kPbWriteBS) * U) {
    typename S::Ids w[U];
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i + u * kPbWriteBS < i1) w[u] = ids[i + u * kPbWriteBS];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i + u * kPbWriteBS >= i1) break;
      uint32_t sid[E];
      S::Unpack(w[u], sid);
      T v[E];
#pragma unroll
      for (int e = 0; e < E; e++) v[e] = slice[sid[e] & m];
      out[i + u * kPbWriteBS] = S::Pack(v);
    }
  }
}

template <class T>
__global__ __launch_bounds__(kPbSumBS) void PbSum(PbArgs<T> a, const PbChunk* __restrict__ chunks) {
  using Word = typename PbAcc<T>::Word;
  __shared__ Word s_acc[1 << kPbMaxLog];
  __shared__ double s_red[kPbSumBS / kWave];
  const PbChunk c = chunks[blockIdx.x];
  const int w = 1 << a.lg;
  const uint32_t m = static_cast<uint32_t>(w - 1);
  for (int i = threadIdx.x; i < w; i += kPbSumBS) s_acc[i] = Word(0);
  __syncthreads();
  for (int64_t q = int64_t(c.q0) + threadIdx.x; q < c.q1; q += kPbSumBS) {
    const uint32_t o = a.oct[q];
    uint32_t d[8];
    Unpack16(a.dst16[o], d);
    T v[8];
    LoadOctet(a.val, o, v);
    PbAddOctet(s_acc, d, v, m);
  }
  __syncthreads();
  const int64_t first = int64_t(c.part) << a.lg;
  const int cnt = static_cast<int>(min(int64_t(w), a.n - first));
  double r = 0.0;
  if (c.slot < 0) {
    for (int i = threadIdx.x; i < cnt; i += kPbSumBS) r += PbFinish(a, first + i, PbValue(s_acc[i]));
  } else {
    Word* p = a.part + (int64_t(c.slot) << a.lg);
    for (int i = threadIdx.x; i < cnt; i += kPbSumBS) p[i] = s_acc[i];
  }
  r = BlockSum<kPbSumBS>(r, s_red);
  if (threadIdx.x == 0 && r != 0.0) atomicAdd(&a.residual[blockIdx.x & (kStripes - 1)], r);
}

template <class T>
__global__ __launch_bounds__(kBS) void PbMerge(PbArgs<T> a, const PbMulti* __restrict__ multi, int32_t n_multi) {
  __shared__ double s_red[kBS / kWave];
  const int64_t w = int64_t(1) << a.lg, total = int64_t(n_multi) << a.lg;
  double r = 0.0;
  for (int64_t j = blockIdx.x * int64_t(kBS) + threadIdx.x; j < total; j += int64_t(gridDim.x) * kBS) {
    const PbMulti M = multi[j >> a.lg];
    const int64_t i = j & (w - 1), u = (int64_t(M.bin) << a.lg) + i;
    if (u >= a.n) continue;
    typename PbAcc<T>::Word sum = a.part[(int64_t(M.slot0) << a.lg) + i];
    for (uint32_t s = 1; s < M.nslots; s++) sum += a.part[(int64_t(M.slot0 + s) << a.lg) + i];  // in slot order
    r += PbFinish(a, u, PbValue(sum));
  }
  AddResidual(a.residual, r, s_red);
}

// Build.  Each block walks in-lists, a wavefront 64 entries at a time, cut
// into runs of equal source tile (sorted lists give one run per tile).
// Work items: the first blocks take a range of one bin's vertices (about
// kPbItemEntries in-entries, cut at multiples of 256 vertices) and walk the
// lists shorter than kPbHubMin, a wavefront per vertex; the blocks after
// them take one chunk (kPbHubChunk entries at most) of a longer list each,
// its wavefronts strided over the chunk.
constexpr int64_t kPbHubMin = int64_t(1) << 15;
constexpr int64_t kPbHubChunk = int64_t(1) << 16;
constexpr int64_t kPbItemEntries = int64_t(1) << 20;

struct PbItem {
  int64_t u0, u1;  // vertices [u0, u1), all in one bin
};

struct PbHubChunk {
  int64_t e0, e1;  // in-entries [e0, e1) of vertex u
  int32_t u, pad;
};

struct PbRun {
  int32_t t;       // the lane's source tile (-1 past the list)
  int first, len;  // its run: first lane, and length (valid in the run's first lane)
  uint64_t starts;
};
__device__ __forceinline__ PbRun PbRunOf(int32_t src, bool valid, int lg, int nvalid) {
  const int lane = LaneId();
  PbRun r;
  r.t = valid ? (src >> lg) : -1;
  const int32_t tp = __shfl_up(r.t, 1);
  const bool start = valid && (lane == 0 || tp != r.t);
  r.starts = __ballot(start);
  const uint64_t rest = lane == kWave - 1 ? 0 : (r.starts >> (lane + 1));
  r.len = start ? (rest ? 1 + __builtin_ctzll(rest) : nvalid - lane) : 0;
  const uint64_t upto = r.starts & ((uint64_t(2) << lane) - 1);
  r.first = upto ? 63 - __builtin_clzll(upto) : 0;
  return r;
}

// Calls f(run, source, valid) for every 64-entry window of this block's work
// item; returns the item's bin.
template <class F>
__device__ __forceinline__ int64_t PbWalk(const int64_t* __restrict__ in_off, const int32_t* __restrict__ in_dst,
                                          int lg, const PbItem* __restrict__ items, int64_t n_items,
                                          const PbHubChunk* __restrict__ hubs, F f) {
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  constexpr int kWaves = kPbBuildBS / kWave;
  auto window = [&](int64_t k0, int64_t e, int32_t u) {
    const bool valid = k0 + lane < e;
    const int32_t s = valid ? in_dst[k0 + lane] : 0;
    f(PbRunOf(s, valid, lg, static_cast<int>(min<int64_t>(kWave, e - k0))), s, valid, u);
  };
  if (blockIdx.x < n_items) {
    const PbItem it = items[blockIdx.x];
    for (int64_t u = it.u0 + wid; u < it.u1; u += kWaves) {
      const int64_t e0 = in_off[u], e = in_off[u + 1];
      if (e - e0 >= kPbHubMin) continue;  // cut into chunks
      for (int64_t k0 = e0; k0 < e; k0 += kWave) window(k0, e, static_cast<int32_t>(u));
    }
    return it.u0 >> lg;
  }
  const PbHubChunk h = hubs[blockIdx.x - n_items];
  for (int64_t k0 = h.e0 + int64_t(wid) * kWave; k0 < h.e1; k0 += int64_t(kWaves) * kWave) window(k0, h.e1, h.u);
  return h.u >> lg;
}

__global__ __launch_bounds__(kPbBuildBS) void PbFindHubs(const int64_t* __restrict__ in_off, int64_t n,
                                                         PbHubChunk* __restrict__ hubs, unsigned* __restrict__ count,
                                                         unsigned cap) {
  for (int64_t u = blockIdx.x * int64_t(kPbBuildBS) + threadIdx.x; u < n; u += int64_t(gridDim.x) * kPbBuildBS) {
    const int64_t e0 = in_off[u], e1 = in_off[u + 1];
    if (e1 - e0 < kPbHubMin) continue;
    const unsigned i = atomicAdd(count, 1u);
    if (i < cap) hubs[i] = PbHubChunk{e0, e1, static_cast<int32_t>(u), 0};
  }
}

// In-entries per (tile, bin) block, added into cnt[t * k + b] (zeroed).
__global__ __launch_bounds__(kPbBuildBS) void PbCount(const int64_t* __restrict__ in_off,
                                                      const int32_t* __restrict__ in_dst, int lg, int64_t k,
                                                      const PbItem* __restrict__ items, int64_t n_items,
                                                      const PbHubChunk* __restrict__ hubs, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t c[kPbMaxTiles];
  for (int64_t t = threadIdx.x; t < k; t += kPbBuildBS) c[t] = 0;
  __syncthreads();
  const int64_t b = PbWalk(in_off, in_dst, lg, items, n_items, hubs, [&](const PbRun& r, int32_t, bool, int32_t) {
    if (r.len > 0) atomicAdd(&c[r.t], static_cast<uint32_t>(r.len));
  });
  __syncthreads();
  for (int64_t t = threadIdx.x; t < k; t += kPbBuildBS)
    if (c[t]) atomicAdd(&cnt[t * k + b], c[t]);
}

// Counts to octets, in place (tile-major) and transposed into bm (bin-major).
__global__ __launch_bounds__(kPbBuildBS) void PbOctets(uint32_t* __restrict__ tm, uint32_t* __restrict__ bm, int64_t k) {
  for (int64_t i = blockIdx.x * int64_t(kPbBuildBS) + threadIdx.x; i < k * k; i += int64_t(gridDim.x) * kPbBuildBS) {
    const uint32_t o = (tm[i] + 7) >> 3;
    tm[i] = o;
    bm[(i % k) * k + i / k] = o;
  }
}

// Writes each entry's offsets into its block: the work item first reserves
// a range per tile from the block's cursor (cur[t * k + b], zeroed), then
// places each run there, so a run's entries (one destination's, within one
// tile) stay together.  run: the exclusive scan of the octet counts.
__global__ __launch_bounds__(kPbBuildBS) void PbEmit(const int64_t* __restrict__ in_off,
                                                     const int32_t* __restrict__ in_dst, int lg, int64_t k,
                                                     const PbItem* __restrict__ items, int64_t n_items,
                                                     const PbHubChunk* __restrict__ hubs, const uint32_t* __restrict__ run,
                                                     uint32_t* __restrict__ cur, uint16_t* __restrict__ src16,
                                                     uint16_t* __restrict__ dst16) {
  __shared__ uint32_t c[kPbMaxTiles], base[kPbMaxTiles];
  for (int64_t t = threadIdx.x; t < k; t += kPbBuildBS) c[t] = 0;
  __syncthreads();
  const int64_t b = PbWalk(in_off, in_dst, lg, items, n_items, hubs, [&](const PbRun& r, int32_t, bool, int32_t) {
    if (r.len > 0) atomicAdd(&c[r.t], static_cast<uint32_t>(r.len));
  });
  __syncthreads();
  for (int64_t t = threadIdx.x; t < k; t += kPbBuildBS) {
    base[t] = c[t] ? atomicAdd(&cur[t * k + b], c[t]) : 0;
    c[t] = 0;
  }
  __syncthreads();
  const uint32_t m = (1u << lg) - 1;
  PbWalk(in_off, in_dst, lg, items, n_items, hubs, [&](const PbRun& r, int32_t s, bool valid, int32_t u) {
    uint32_t at = 0;
    if (r.len > 0) at = base[r.t] + atomicAdd(&c[r.t], static_cast<uint32_t>(r.len));
    at = __shfl(at, r.first);
    if (valid) {
      const int64_t pos = int64_t(run[int64_t(r.t) * k + b]) * 8 + at + (LaneId() - r.first);
      src16[pos] = static_cast<uint16_t>(static_cast<uint32_t>(s) & m);
      dst16[pos] = static_cast<uint16_t>(static_cast<uint32_t>(u) & m);
    }
  });
}

// dst[i] = src[i * stride] for i < count (strided values for the host).
template <class T>
__global__ __launch_bounds__(kPbBuildBS) void PbGather(const T* __restrict__ src, int64_t stride, int64_t count,
                                                       T* __restrict__ dst) {
  for (int64_t i = blockIdx.x * int64_t(kPbBuildBS) + threadIdx.x; i < count; i += int64_t(gridDim.x) * kPbBuildBS)
    dst[i] = src[i * stride];
}

// Pads each block from its entries to its octets with (0, kPbPad).
__global__ __launch_bounds__(kPbBuildBS) void PbPad(const uint32_t* __restrict__ run, const uint32_t* __restrict__ cur,
                                                    int64_t blocks, uint16_t* __restrict__ src16,
                                                    uint16_t* __restrict__ dst16) {
  for (int64_t i = blockIdx.x * int64_t(kPbBuildBS) + threadIdx.x; i < blocks; i += int64_t(gridDim.x) * kPbBuildBS) {
    const int64_t p1 = int64_t(run[i + 1]) * 8;
    for (int64_t p = int64_t(run[i]) * 8 + cur[i]; p < p1; p++) {
      src16[p] = 0;
      dst16[p] = static_cast<uint16_t>(kPbPad);
    }
  }
}

// The bin-major octet index: bin b lists the octets of blocks (0, b),
// (1, b), ... in turn.  bpos: the exclusive scan of bm.
__global__ __launch_bounds__(kPbBuildBS) void PbOctIndex(int64_t k, const uint32_t* __restrict__ run,
                                                         const uint32_t* __restrict__ bpos, uint32_t* __restrict__ oct) {
  const int64_t b = blockIdx.x;
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  for (int64_t t = wid; t < k; t += kPbBuildBS / kWave) {
    const uint32_t r0 = run[t * k + b], len = run[t * k + b + 1] - r0, d0 = bpos[b * k + t];
    for (uint32_t i = lane; i < len; i += kWave) oct[d0 + i] = r0 + i;
  }
}

// The width the Jacobi run of this precision uses (0: no blocks).  By
// default about 4096 tiles: 2^(ceil(log2 n) - 12) vertices, from 2^10 to
// 2^14 (measured on Kronecker graphs of 2^20 to 2^26 vertices: the
// narrowest width with few enough tiles to keep the padding and the build
// tables small, since narrower blocks give more bins, so fewer bins summed
// in several chunks, and smaller LDS footprints).
int BlockWidth(const gg_graph* g, bool fp64) {
  const int knob = g->tune.pr_blocks_log2;
  if (knob > 0) return knob;
  if (knob < 0 || g->tune.pr_segment_log2[fp64 ? 1 : 0] != 0) return 0;
  const int64_t n = g->n;
  if (n < (int64_t(1) << 20) || n > (int64_t(1) << 27)) return 0;
  const int ceil_log2 = 64 - __builtin_clzll(static_cast<unsigned long long>(n - 1));
  return std::min(kPbMaxLog, std::max(10, ceil_log2 - 12));
}

// The propagation-block layout, built on first use (and again when the
// width changes); null when the Jacobi run of this precision goes without.
const gg_graph::PrBlocks* Blocks(gg_graph* g, bool fp64) {
  const int lg = BlockWidth(g, fp64);
  if (lg == 0) return nullptr;
  gg_graph::PrBlocks& B = g->pr_blocks;
  if (B.built && B.lg == lg) return &B;
  const int64_t n = g->n, k = (n + (int64_t(1) << lg) - 1) >> lg;
  GG_REQUIRE(k <= kPbMaxTiles, GG_EINVAL, "PageRank block width too small for this graph (more than 8192 tiles)");
  // Octet positions are 32-bit: at most 7 padding entries per block.
  GG_REQUIRE((g->in_entries + 7 * k * k) / 8 < (int64_t(1) << 32), GG_EINVAL, "too many in-entries for PageRank blocks");
  g->ws.Drop("pr.pb.");
  WallTimer wt;
  B = gg_graph::PrBlocks{};
  const bool prof = g->profile;
  g->profile = false;  // the build is not part of any measured call
  struct Restore {
    gg_graph* g;
    bool prof;
    ~Restore() { g->profile = prof; }
  } restore{g, prof};
  hipStream_t st = g->stream;
  const int64_t kk = k * k;
  // Lists of kPbHubMin entries or more, cut into chunks (their own work items).
  const unsigned hub_cap = static_cast<unsigned>(g->in_entries / kPbHubMin + 1);
  PbHubChunk* hubs = g->Scratch<PbHubChunk>("pr.pbb.hubs", hub_cap);
  unsigned* nh = g->Scratch<unsigned>("pr.pbb.nh", 1);
  GG_HIP(hipMemsetAsync(nh, 0, sizeof(unsigned), st));
  PbFindHubs<<<GridFor(n, kPbBuildBS, g->cu_count * 8), kPbBuildBS, 0, st>>>(g->in_off, n, hubs, nh, hub_cap);
  unsigned nhubs = 0;
  GG_HIP(hipMemcpyAsync(&nhubs, nh, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  GG_REQUIRE(nhubs <= hub_cap, GG_EHIP, "PageRank block build: hub list overflow");
  std::vector<PbHubChunk> hl(nhubs), chunks;
  if (nhubs) GG_HIP(hipMemcpy(hl.data(), hubs, nhubs * sizeof(PbHubChunk), hipMemcpyDeviceToHost));
  std::sort(hl.begin(), hl.end(), [](const PbHubChunk& x, const PbHubChunk& y) { return x.u < y.u; });
  for (const PbHubChunk& h : hl)
    for (int64_t e = h.e0; e < h.e1; e += kPbHubChunk)
      chunks.push_back(PbHubChunk{e, std::min(e + kPbHubChunk, h.e1), h.u, 0});
  // Vertex-range items: each bin cut at multiples of G vertices once its
  // short lists pass kPbItemEntries entries (offsets sampled every G).
  const int64_t G = std::min<int64_t>(256, int64_t(1) << lg), ng = (n + G - 1) / G;
  std::vector<int64_t> goff(ng + 1), hub_in(ng, 0);
  int64_t* dgoff = g->Scratch<int64_t>("pr.pbb.goff", ng);
  PbGather<int64_t><<<GridFor(ng, kPbBuildBS, g->cu_count * 4), kPbBuildBS, 0, st>>>(g->in_off, G, ng, dgoff);
  GG_HIP(hipMemcpyAsync(goff.data(), dgoff, ng * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GG_HIP(hipMemcpyAsync(&goff[ng], g->in_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  for (const PbHubChunk& h : hl) hub_in[h.u / G] += h.e1 - h.e0;
  std::vector<PbItem> il;
  for (int64_t b = 0; b < k; b++) {
    const int64_t g0 = (b << lg) / G, g1 = std::min(ng, ((b + 1) << lg) / G);
    int64_t start = g0, acc = 0;
    for (int64_t gi = g0; gi < g1; gi++) {
      const int64_t e = goff[gi + 1] - goff[gi] - hub_in[gi];
      if (acc > 0 && acc + e > kPbItemEntries) {
        il.push_back(PbItem{start * G, gi * G});
        start = gi;
        acc = 0;
      }
      acc += e;
    }
    il.push_back(PbItem{start * G, std::min(n, g1 * G)});
  }
  const int64_t n_items = static_cast<int64_t>(il.size());
  PbItem* ditems = g->Scratch<PbItem>("pr.pbb.items", il.size());
  GG_HIP(hipMemcpyAsync(ditems, il.data(), il.size() * sizeof(PbItem), hipMemcpyHostToDevice, st));
  PbHubChunk* dchunks = g->Scratch<PbHubChunk>("pr.pbb.chunks", chunks.size() + 1);
  if (!chunks.empty())
    GG_HIP(hipMemcpyAsync(dchunks, chunks.data(), chunks.size() * sizeof(PbHubChunk), hipMemcpyHostToDevice, st));
  const unsigned items = static_cast<unsigned>(n_items + static_cast<int64_t>(chunks.size()));
  uint32_t* tm = g->Scratch<uint32_t>("pr.pbb.tm", kk + 1);
  uint32_t* bm = g->Scratch<uint32_t>("pr.pbb.bm", kk + 1);
  uint32_t* cur = g->Scratch<uint32_t>("pr.pbb.cur", kk);
  GG_HIP(hipMemsetAsync(tm, 0, kk * sizeof(uint32_t), st));
  GG_HIP(hipMemsetAsync(cur, 0, kk * sizeof(uint32_t), st));
  PbCount<<<items, kPbBuildBS, 0, st>>>(g->in_off, g->in_dst, lg, k, ditems, n_items, dchunks, tm);
  PbOctets<<<GridFor(kk, kPbBuildBS, g->cu_count * 8), kPbBuildBS, 0, st>>>(tm, bm, k);
  GG_HIP(hipGetLastError());
  DeviceScan<uint32_t>(g, "pr.pbb.st.", tm, kk, tm + kk);
  DeviceScan<uint32_t>(g, "pr.pbb.sb.", bm, kk, bm + kk);
  // Tile starts tm[t * k] and bin starts bm[b * k], t, b = 0..k (the last
  // one is the total).
  std::vector<uint32_t> tile_at(k + 1), bin_at(k + 1);
  uint32_t* starts = g->Scratch<uint32_t>("pr.pbb.starts", 2 * (k + 1));
  PbGather<uint32_t><<<GridFor(k + 1, kPbBuildBS), kPbBuildBS, 0, st>>>(tm, k, k + 1, starts);
  PbGather<uint32_t><<<GridFor(k + 1, kPbBuildBS), kPbBuildBS, 0, st>>>(bm, k, k + 1, starts + k + 1);
  GG_HIP(hipMemcpyAsync(tile_at.data(), starts, (k + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  GG_HIP(hipMemcpyAsync(bin_at.data(), starts + k + 1, (k + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  const int64_t octets = tile_at[k];
  GG_REQUIRE(bin_at[k] == tile_at[k], GG_EHIP, "PageRank block tables disagree");
  uint16_t* src16 = g->Scratch<uint16_t>("pr.pb.src", 8 * octets + 8);
  uint16_t* dst16 = g->Scratch<uint16_t>("pr.pb.dst", 8 * octets + 8);
  uint32_t* oct = g->Scratch<uint32_t>("pr.pb.oct", octets + 1);
  PbEmit<<<items, kPbBuildBS, 0, st>>>(g->in_off, g->in_dst, lg, k, ditems, n_items, dchunks, tm, cur, src16, dst16);
  PbPad<<<GridFor(kk, kPbBuildBS, g->cu_count * 8), kPbBuildBS, 0, st>>>(tm, cur, kk, src16, dst16);
  PbOctIndex<<<static_cast<unsigned>(k), kPbBuildBS, 0, st>>>(k, tm, bm, oct);
  GG_HIP(hipGetLastError());
  // Chunks: the write phase's per tile (at most 4 * 2^lg octets), the sum
  // phase's per bin (at most 16 * 2^lg octets, so fewer bins need partial
  // sums and a merge; every bin has one, so every vertex is finished).  On
  // smaller graphs the caps shrink to give each phase about four chunks per
  // CU (down to 2^12 octets a chunk).
  const int64_t per_cu = std::max<int64_t>(int64_t(1) << 12, octets / (4 * int64_t(g->cu_count)));
  const int64_t cap = std::min(int64_t(4) << lg, per_cu), scap = std::min(int64_t(16) << lg, per_cu);
  std::vector<PbChunk> wc, sc;
  std::vector<PbMulti> multi;
  int64_t slots = 0;
  for (int64_t t = 0; t < k; t++)
    for (int64_t q = tile_at[t]; q < tile_at[t + 1]; q += cap)
      wc.push_back(PbChunk{static_cast<uint32_t>(t), -1, static_cast<uint32_t>(q),
                           static_cast<uint32_t>(std::min<int64_t>(q + cap, tile_at[t + 1]))});
  for (int64_t b = 0; b < k; b++) {
    const int64_t q0 = bin_at[b], q1 = bin_at[b + 1];
    const int64_t pieces = std::max<int64_t>(1, (q1 - q0 + scap - 1) / scap);
    if (pieces == 1) {
      sc.push_back(PbChunk{static_cast<uint32_t>(b), -1, static_cast<uint32_t>(q0), static_cast<uint32_t>(q1)});
      continue;
    }
    multi.push_back(PbMulti{static_cast<uint32_t>(b), static_cast<uint32_t>(slots), static_cast<uint32_t>(pieces), 0});
    for (int64_t i = 0; i < pieces; i++)
      sc.push_back(PbChunk{static_cast<uint32_t>(b), static_cast<int32_t>(slots + i), static_cast<uint32_t>(q0 + i * scap),
                           static_cast<uint32_t>(std::min(q0 + (i + 1) * scap, q1))});
    slots += pieces;
  }
  auto upload = [&](const char* name, const void* p, size_t bytes) {
    void* d = g->ws.Get(name, std::max<size_t>(bytes, 16));
    if (bytes) GG_HIP(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, st));
  };
  upload("pr.pb.wchunks", wc.data(), wc.size() * sizeof(PbChunk));
  upload("pr.pb.schunks", sc.data(), sc.size() * sizeof(PbChunk));
  upload("pr.pb.multi", multi.data(), multi.size() * sizeof(PbMulti));
  GG_HIP(hipStreamSynchronize(st));
  g->ws.Drop("pr.pbb.");
  B.built = true;
  B.lg = lg;
  B.k = k;
  B.octets = octets;
  B.n_write = static_cast<int64_t>(wc.size());
  B.n_sum = static_cast<int64_t>(sc.size());
  B.n_multi = static_cast<int64_t>(multi.size());
  B.slots = slots;
  B.build_ms = wt.ms();
  return &B;
}

// in_place: one contribution array, each vertex's contribution replaced as
// soon as its score is (PageRankPullGS, pr.cc:34-61), and each sweep runs as
// kGsPasses launches over consecutive vertex-id ranges, so later ranges
// read the new contributions of earlier ones -- the reference's threads
// sweep ids in order within their 16384-vertex chunks.  (Ordering the
// passes by degree bin instead converges more slowly than Jacobi on
// Kronecker graphs.)
template <class T>
void RunPageRank(gg_graph* g, int max_iters, double epsilon, bool in_place, float* scores_out, gg_stats* stats,
                 const WallTimer& wt) {
  using Stored = typename Acc<T>::Stored;
  const bool fp64 = sizeof(T) == 8;  // profile names carry the precision
  const int64_t n = g->n;
  const PrBins bins = GetBins(g);
  const gg_graph::PrBlocks* pb = in_place ? nullptr : Blocks(g, fp64);
  const gg_graph::PrSegments* seg = in_place || pb ? nullptr : Segments<T>(g);
  const std::string seg_tag = sizeof(T) == 4 ? "pr.seg32." : "pr.seg64.";
  const int32_t* seg_src = seg ? g->Scratch<int32_t>((seg_tag + "src").c_str(), seg->hot_entries) : nullptr;
  const int32_t* seg_dst = seg ? g->Scratch<int32_t>((seg_tag + "dst").c_str(), seg->pairs) : nullptr;
  const int64_t* seg_off = seg ? g->Scratch<int64_t>((seg_tag + "off").c_str(), seg->pairs + 1) : nullptr;
  const int32_t* seg_tail = seg ? g->Scratch<int32_t>((seg_tag + "tail").c_str(), n) : nullptr;
  Stored* acc = seg ? g->Scratch<Stored>(sizeof(T) == 4 ? "pr.acc32" : "pr.acc64", n) : nullptr;
  const int width = g->tune.pr_blocks_per_cu;  // gather blocks per CU
  T* score = g->Scratch<T>(sizeof(T) == 4 ? "pr.score32" : "pr.score64", n);
  T* ca = g->Scratch<T>(sizeof(T) == 4 ? "pr.contrib32a" : "pr.contrib64a", n);
  T* cb = g->Scratch<T>(sizeof(T) == 4 ? "pr.contrib32b" : "pr.contrib64b", n);
  double* residual = g->Scratch<double>("pr.residual", kStripes);
  // Constants exactly as the reference computes them, in float
  // (pr_spmv.cc:32, 38-39), then widened for the double variant.
  const float kDampF = 0.85f;
  const float init_f = 1.0f / static_cast<float>(n);
  const float base_f = (1.0f - kDampF) / static_cast<float>(n);
  const T init = static_cast<T>(init_f), base = static_cast<T>(base_f), damp = static_cast<T>(kDampF);
  const double N = static_cast<double>(n), E = static_cast<double>(g->in_entries);
  const double vbytes = sizeof(T) == 4 ? 16.0 : 32.0;  // contrib read/write + score read/write per vertex
  const dim3 blk(kBS);
  int64_t launches = 0;

  BeginTiming(g);
  Run(g, "pr_reset", 0, PrResetResidual, dim3(1), blk, residual);
  Run(g, fp64 ? "pr_init_fp64" : "pr_init", 8 * (N + 1) + 2 * sizeof(T) * N, PrInit<T>, dim3(GridFor(n, kBS, g->cu_count * 32)),
      blk, static_cast<const int64_t*>(g->out_off), score, ca, acc, n, init);
  launches += 2;
  int iters = 0;
  double error = 0.0;
  // Propagation blocks: per iteration a write, a sum and (for bins summed
  // in several chunks) a merge launch.  Credited bytes: the write reads
  // 2 B of offsets and writes a value per padded entry, plus the tile
  // contributions; the sum reads 4 B of index per octet, 2 B of offsets and
  // a value per entry, plus each vertex's score, offsets and new values
  // (or its partial sums); the merge reads the partial sums.
  T* pb_val = pb ? g->Scratch<T>(fp64 ? "pr.pb.val64" : "pr.pb.val32", 8 * pb->octets + 8) : nullptr;
  using Word = typename PbAcc<T>::Word;
  Word* pb_part = pb && pb->slots ? g->Scratch<Word>(fp64 ? "pr.pb.part64" : "pr.pb.part32", pb->slots << pb->lg)
                                  : nullptr;
  for (int it = 0; pb && it < max_iters; it++) {
    PbArgs<T> a{ca, pb_val,
                static_cast<const uint4*>(g->ws.Get("pr.pb.src", 0)), static_cast<const uint4*>(g->ws.Get("pr.pb.dst", 0)),
                static_cast<const uint32_t*>(g->ws.Get("pr.pb.oct", 0)), pb_part, g->out_off, score, cb, base, damp,
                residual, n, pb->lg};
    const double ent = 8.0 * pb->octets, vs = sizeof(T), ps = sizeof(Word);
    if (pb->n_write)
      Run(g, fp64 ? "pr_blocks_write_fp64" : "pr_blocks_write", ent * (2 + vs) + vs * N, PbWrite<T>,
          dim3(static_cast<unsigned>(pb->n_write)), dim3(kPbWriteBS), a,
          static_cast<const PbChunk*>(g->ws.Get("pr.pb.wchunks", 0)));
    Run(g, fp64 ? "pr_blocks_sum_fp64" : "pr_blocks_sum",
        ent * (2 + vs) + 0.5 * ent + N * (8 + 3 * vs) + ps * double(pb->slots << pb->lg), PbSum<T>,
        dim3(static_cast<unsigned>(pb->n_sum)), dim3(kPbSumBS), a, static_cast<const PbChunk*>(g->ws.Get("pr.pb.schunks", 0)));
    launches += pb->n_write ? 2 : 1;
    if (pb->n_multi) {
      Run(g, fp64 ? "pr_blocks_merge_fp64" : "pr_blocks_merge", ps * double(pb->slots << pb->lg), PbMerge<T>,
          dim3(GridFor(pb->n_multi << pb->lg, kBS, g->cu_count * 16)), blk, a,
          static_cast<const PbMulti*>(g->ws.Get("pr.pb.multi", 0)), static_cast<int32_t>(pb->n_multi));
      launches++;
    }
    const unsigned long long seq = MailNext(g);
    Run(g, "pr_publish", 0, PrPublish, dim3(1), blk, residual, g->mail_dev, seq);
    launches++;
    unsigned long long bits = 0;
    MailWait(g, seq, &bits, 1);
    std::memcpy(&error, &bits, sizeof(error));
    std::swap(ca, cb);
    iters++;
    if (error < epsilon) break;
  }
  for (int it = 0; !pb && it < max_iters; it++) {
    PrArgs<T> a{g->in_off, g->in_dst, g->out_off, ca, in_place ? ca : cb, score, base, damp, residual, 0, 0,
                seg_tail, acc};
    const int passes = in_place ? kGsPasses : 1;
    // Algorithmic bytes of one iteration: offsets 8 B and contrib/score
    // traffic per vertex, 4 B per in-neighbour id (each array counted once;
    // the random contribution gathers are not).  A Gauss-Seidel pass is
    // credited its share; with segments, each segment launch is credited
    // its entries' ids and the finishing gather the rest.  One gather launch
    // per pass: the three bins' blocks run side by side.
    const double per_vertex = 8.0 + vbytes, share = 1.0 / passes;
    double bytes = per_vertex * n + 4.0 * (bins.e_small + bins.e_medium + bins.e_large);
    if (seg) {
      for (int s = 0; s < seg->ns; s++) {
        const int64_t p0 = seg->first_pair[s];
        const int64_t nl = seg->large[s], nm = seg->medium[s], nsm = seg->first_pair[s + 1] - p0 - nl - nm;
        if (nl + nm + nsm == 0) continue;
        const int32_t gl = static_cast<int32_t>(nl ? GridFor(nl, 1, g->cu_count * (width + 1) / 2) : 0);
        const int32_t gm = static_cast<int32_t>(nm ? GridFor(nm, kBS / kWave, g->cu_count * width) : 0);
        const int32_t gs = static_cast<int32_t>(nsm ? GridFor(nsm, kBS, g->cu_count * width) : 0);
        SegArgs<T> sa{ca, seg_src, seg_dst, seg_off, acc, p0, static_cast<int32_t>(nl), static_cast<int32_t>(nm),
                      static_cast<int32_t>(nsm), gl, gm};
        if (s < seg->n_lds) {
          const int64_t lo = int64_t(s) << seg->sbl;
          const int32_t cnt = static_cast<int32_t>(std::min<int64_t>(int64_t(1) << seg->sbl, n - lo));
          Run(g, fp64 ? "pr_segment_lds_fp64" : "pr_segment_lds", 4.0 * seg->entries[s], PrSegmentLds<T>, dim3(g->cu_count), dim3(kLdsBS), sa, lo,
              cnt);
        } else {
          Run(g, fp64 ? "pr_segment_fp64" : "pr_segment", 4.0 * seg->entries[s], PrSegment<T>, dim3(gl + gm + gs), blk, sa);
        }
        launches++;
      }
      bytes -= 4.0 * seg->hot_entries;
    }
    const int32_t nb_large =
        static_cast<int32_t>(bins.n_large ? GridFor(bins.n_large, 1, g->cu_count * (width + 1) / 2) : 0);
    const int32_t nb_medium =
        static_cast<int32_t>(bins.n_medium ? GridFor(bins.n_medium, kBS / kWave, g->cu_count * width) : 0);
    const int32_t nb_small =
        static_cast<int32_t>(bins.n_small ? GridFor(bins.n_small, kBS, g->cu_count * width) : 0);
    for (int pass = 0; pass < passes; pass++) {
      a.lo = static_cast<int32_t>(n * pass / passes);
      a.hi = static_cast<int32_t>(n * (pass + 1) / passes);
      Run(g, fp64 ? "pr_gather_fp64" : "pr_gather", share * bytes, PrGather<T>, dim3(nb_large + nb_medium + nb_small), blk, a,
          static_cast<const int32_t*>(bins.small), static_cast<int32_t>(bins.n_small),
          static_cast<const int32_t*>(bins.medium), static_cast<int32_t>(bins.n_medium),
          static_cast<const int32_t*>(bins.large), static_cast<int32_t>(bins.n_large), nb_large, nb_medium);
      launches++;
    }
    const unsigned long long seq = MailNext(g);
    Run(g, "pr_publish", 0, PrPublish, dim3(1), blk, residual, g->mail_dev, seq);
    launches++;
    unsigned long long bits = 0;
    MailWait(g, seq, &bits, 1);
    std::memcpy(&error, &bits, sizeof(error));
    if (!in_place) std::swap(ca, cb);
    iters++;
    if (error < epsilon) break;
  }
  float* out32 = nullptr;
  if (sizeof(T) == 8) {
    out32 = g->Scratch<float>("pr.out32", n);
    Run(g, "pr_to_float", 12 * N, ToFloat, dim3(GridFor(n, kBS, g->cu_count * 32)), blk,
        reinterpret_cast<const double*>(score), out32, n);
    launches++;
  } else {
    out32 = reinterpret_cast<float*>(score);
  }
  GG_HIP(hipGetLastError());
  FinishTiming(g, stats, wt);
  if (scores_out) CopyOutVertexF32(g, out32, scores_out);
  if (stats) {
    stats->total_ms = wt.ms();
    stats->iterations = iters;
    stats->residual = error;
    stats->edges_examined = static_cast<int64_t>(E) * iters;
    stats->launches = launches;
  }
}

}  // namespace

const gg_graph::PrBlocks* PropagationBlocks(gg_graph* g) { return Blocks(g, false); }

PbArrays PropagationBlockArrays(gg_graph* g) {
  return PbArrays{static_cast<const uint4*>(g->ws.Get("pr.pb.src", 0)),
                  static_cast<const uint4*>(g->ws.Get("pr.pb.dst", 0)),
                  static_cast<const uint32_t*>(g->ws.Get("pr.pb.oct", 0)),
                  static_cast<const PbChunk*>(g->ws.Get("pr.pb.wchunks", 0)),
                  static_cast<const PbChunk*>(g->ws.Get("pr.pb.schunks", 0)),
                  static_cast<const PbMulti*>(g->ws.Get("pr.pb.multi", 0))};
}

}  // namespace gg

using namespace gg;

extern "C" int gg_pagerank(gg_graph* g, int max_iters, double epsilon, int mode, float* scores,
                           gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "PageRank needs an unweighted graph upload");
    GG_REQUIRE(g->n > 0, GG_EINVAL, "empty graph");
    GG_REQUIRE(max_iters >= 0, GG_EINVAL, "max_iters must be >= 0");
    GG_REQUIRE(mode == GG_PR_FP32 || mode == GG_PR_FP64 || mode == GG_PR_GS, GG_EINVAL, "unknown mode");
    GG_HIP(hipSetDevice(g->device));
    if (mode == GG_PR_FP64)
      RunPageRank<double>(g, max_iters, epsilon, false, scores, stats, wt);
    else
      RunPageRank<float>(g, max_iters, epsilon, mode == GG_PR_GS, scores, stats, wt);
  });
}

extern "C" int gg_pagerank_layout(gg_graph* g, int mode, int64_t* info) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr && info != nullptr, GG_EINVAL, "null argument");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "PageRank needs an unweighted graph upload");
    GG_REQUIRE(g->n > 0, GG_EINVAL, "empty graph");
    GG_REQUIRE(mode == GG_PR_FP32 || mode == GG_PR_FP64, GG_EINVAL, "a Jacobi mode");
    GG_HIP(hipSetDevice(g->device));
    for (int i = 0; i < 6; i++) info[i] = 0;
    if (BlockWidth(g, mode == GG_PR_FP64) > 0) return;  // the run goes by propagation blocks
    const gg_graph::PrSegments* L = mode == GG_PR_FP64 ? Segments<double>(g) : Segments<float>(g);
    if (!L) return;
    info[0] = L->n_lds;
    info[1] = L->ns - L->n_lds;
    info[2] = L->pairs;
    info[3] = L->hot_entries;
    info[4] = L->hot;
    info[5] = static_cast<int64_t>(L->build_ms * 1e3);
  });
}

extern "C" int gg_pagerank_blocks(gg_graph* g, int mode, int64_t* info) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr && info != nullptr, GG_EINVAL, "null argument");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "PageRank needs an unweighted graph upload");
    GG_REQUIRE(g->n > 0, GG_EINVAL, "empty graph");
    GG_REQUIRE(mode == GG_PR_FP32 || mode == GG_PR_FP64, GG_EINVAL, "a Jacobi mode");
    GG_HIP(hipSetDevice(g->device));
    for (int i = 0; i < 6; i++) info[i] = 0;
    const gg_graph::PrBlocks* B = Blocks(g, mode == GG_PR_FP64);
    if (!B) return;
    info[0] = B->lg;
    info[1] = B->k;
    info[2] = 8 * B->octets;
    info[3] = B->n_write;
    info[4] = B->n_sum;
    info[5] = static_cast<int64_t>(B->build_ms * 1e3);
  });
}
