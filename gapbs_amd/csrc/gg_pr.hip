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
// iteration is one gather launch per bin.  Each gather writes the vertex's
// new score and, for the next iteration, its new contribution into a second
// buffer, so the reference's separate contrib pass is fused away.  Two
// precisions:
//   GG_PR_FP32  float scores and contributions, each vertex's sum kept as a
//               compensated (Neumaier) pair and merged across lanes with
//               TwoSum, so the sum carries about twice float's precision;
//   GG_PR_FP64  double scores, contributions and sums.
// The convergence residual is always accumulated in double.
#include "gg_common.h"

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

template <class T>
struct Acc;
template <>
struct Acc<float> {
  using Type = CompSum;
};
template <>
struct Acc<double> {
  using Type = PlainSum;
};

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
};

// Adds contrib[in_dst[k]] for k = k0, k0 + step, ... < e to acc, with
// four independent gathers in flight per lane.
template <class T, class A>
__device__ __forceinline__ void GatherStrided(const PrArgs<T>& a, A& acc, int64_t k0, int64_t e, int step) {
  int64_t k = k0;
  for (; k + 3 * step < e; k += 4 * step) {
    const int32_t v0 = a.in_dst[k], v1 = a.in_dst[k + step], v2 = a.in_dst[k + 2 * step],
                  v3 = a.in_dst[k + 3 * step];
    const T x0 = a.contrib_in[v0], x1 = a.contrib_in[v1], x2 = a.contrib_in[v2], x3 = a.contrib_in[v3];
    acc.Add(x0);
    acc.Add(x1);
    acc.Add(x2);
    acc.Add(x3);
  }
  for (; k < e; k += step) acc.Add(a.contrib_in[a.in_dst[k]]);
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
    const int64_t b = a.in_off[u], e = a.in_off[u + 1];
    A acc = A::Zero();
    int64_t k = b;
    for (; k + 4 <= e; k += 4) {  // four independent gathers in flight
      const int32_t v0 = a.in_dst[k], v1 = a.in_dst[k + 1], v2 = a.in_dst[k + 2], v3 = a.in_dst[k + 3];
      const T x0 = a.contrib_in[v0], x1 = a.contrib_in[v1], x2 = a.contrib_in[v2], x3 = a.contrib_in[v3];
      acc.Add(x0);
      acc.Add(x1);
      acc.Add(x2);
      acc.Add(x3);
    }
    for (; k < e; k++) acc.Add(a.contrib_in[a.in_dst[k]]);
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
    const int64_t b = a.in_off[u], e = a.in_off[u + 1];
    A acc = A::Zero();
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
    const int64_t b = a.in_off[u], e = a.in_off[u + 1];
    A acc = A::Zero();
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
                                              T* __restrict__ contrib, int64_t n, T init) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    score[i] = init;
    contrib[i] = init / static_cast<T>(out_off[i + 1] - out_off[i]);
  }
}

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
  const int64_t n = g->n;
  const PrBins bins = GetBins(g);
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
  Run(g, "pr_init", 8 * (N + 1) + 2 * sizeof(T) * N, PrInit<T>, dim3(GridFor(n, kBS, g->cu_count * 32)),
      blk, static_cast<const int64_t*>(g->out_off), score, ca, n, init);
  launches += 2;
  int iters = 0;
  double error = 0.0;
  for (int it = 0; it < max_iters; it++) {
    PrArgs<T> a{g->in_off, g->in_dst, g->out_off, ca, in_place ? ca : cb, score, base, damp, residual, 0, 0};
    const int passes = in_place ? kGsPasses : 1;
    // Algorithmic bytes of one iteration: offsets 8 B and contrib/score
    // traffic per vertex, 4 B per in-neighbour id (each array counted once;
    // the random contribution gathers are not).  A Gauss-Seidel pass is
    // credited its share.  One launch per pass: the three bins' blocks run
    // side by side.
    const double per_vertex = 8.0 + vbytes, share = 1.0 / passes;
    const double bytes = per_vertex * n + 4.0 * (bins.e_small + bins.e_medium + bins.e_large);
    const int32_t nb_large =
        static_cast<int32_t>(bins.n_large ? GridFor(bins.n_large, 1, g->cu_count * (width + 1) / 2) : 0);
    const int32_t nb_medium =
        static_cast<int32_t>(bins.n_medium ? GridFor(bins.n_medium, kBS / kWave, g->cu_count * width) : 0);
    const int32_t nb_small =
        static_cast<int32_t>(bins.n_small ? GridFor(bins.n_small, kBS, g->cu_count * width) : 0);
    for (int pass = 0; pass < passes; pass++) {
      a.lo = static_cast<int32_t>(n * pass / passes);
      a.hi = static_cast<int32_t>(n * (pass + 1) / passes);
      Run(g, "pr_gather", share * bytes, PrGather<T>, dim3(nb_large + nb_medium + nb_small), blk, a,
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
