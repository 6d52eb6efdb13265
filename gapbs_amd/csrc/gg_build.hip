// gg_graph_build: the squished CSR of an edge list, built on the device;
// gg_graph_generate: the generator's edge list made on the device first.
//
// Reference: BuilderBase::MakeGraphFromEL, MakeCSR and SquishCSR
// (src/builder.h:300-341, 141-186).  The result is byte-identical to the
// host builder's (gh_build_edges) for the same edge list:
//   * vertex count = largest endpoint + 1 (FindMaxNodeID, builder.h:69-78);
//   * a symmetrised graph holds both directions of every edge with its
//     weight; a directed one gets its out-lists and its in-lists (the
//     transpose, builder.h:172-186);
//   * every list is sorted by (destination, weight), duplicates are dropped
//     keeping the first -- the smallest weight -- and self-loops removed.
//
// Every list entry becomes one 64-bit key
//   row << (id_bits + w_bits) | destination << w_bits | (weight - min weight)
// and a stable LSD radix sort of the keys (gg_sort.h) puts each row's entries
// in (destination, weight) order.  One pass over the sorted keys then keeps
// the first entry of each (row, destination) run that is not a self-loop
// (its output slot comes from a scan of the keep flags) and records where
// each row starts and ends; a scan of those row sizes gives the offsets.
// Every step is a streaming pass over HBM.  The temporaries -- the edge list
// in HBM (12 bytes per edge with weights) and, per list entry before the
// squish (2 per edge when symmetrised), two 8-byte keys and a 4-byte flag --
// are freed before the call returns; the lists keep room for every
// unsquished entry (the squish only drops duplicates and self-loops).
#include "gg_common.h"
#include "gg_sort.h"

namespace gg {
namespace {

constexpr int kBS = 256;

int BitsFor(uint64_t max_value) {
  int b = 0;
  while (b < 64 && (max_value >> b) != 0) b++;
  return b;
}

struct Range {
  int lo_id, hi_id, lo_w, hi_w;
};

template <class Op>
__device__ __forceinline__ int WaveReduce(int x, Op op) {
  for (int o = 32; o > 0; o >>= 1) x = op(x, __shfl_xor(x, o));
  return x;
}

// Smallest and largest endpoint and weight.
__global__ __launch_bounds__(kBS) void BdRange(const int32_t* __restrict__ u, const int32_t* __restrict__ v,
                                               const int32_t* __restrict__ w, int64_t m, Range* r) {
  int lo = INT32_MAX, hi = INT32_MIN, wlo = INT32_MAX, whi = INT32_MIN;
  for (int64_t e = blockIdx.x * int64_t(kBS) + threadIdx.x; e < m; e += int64_t(gridDim.x) * kBS) {
    const int32_t a = u[e], b = v[e];
    lo = min(lo, min(a, b));
    hi = max(hi, max(a, b));
    if (w) {
      wlo = min(wlo, w[e]);
      whi = max(whi, w[e]);
    }
  }
  auto mn = [](int a, int b) { return a < b ? a : b; };
  auto mx = [](int a, int b) { return a > b ? a : b; };
  lo = WaveReduce(lo, mn);
  hi = WaveReduce(hi, mx);
  wlo = WaveReduce(wlo, mn);
  whi = WaveReduce(whi, mx);
  if (LaneId() == 0) {
    atomicMin(&r->lo_id, lo);
    atomicMax(&r->hi_id, hi);
    atomicMin(&r->lo_w, wlo);
    atomicMax(&r->hi_w, whi);
  }
}

struct KeyShape {
  int id_bits, w_bits;
  int32_t w_lo;
};

__device__ __forceinline__ uint64_t Pack(uint32_t row, uint32_t dest, uint32_t x, KeyShape s) {
  return (uint64_t(row) << (s.id_bits + s.w_bits)) | (uint64_t(dest) << s.w_bits) | x;
}

// Keys of the list entries: which = 0 (u -> v), 1 (v -> u) or 2 (both, the
// second copy at m + e).
__global__ __launch_bounds__(kBS) void BdKeys(const int32_t* __restrict__ u, const int32_t* __restrict__ v,
                                              const int32_t* __restrict__ w, int64_t m, KeyShape s, int which,
                                              uint64_t* __restrict__ keys) {
  for (int64_t e = blockIdx.x * int64_t(kBS) + threadIdx.x; e < m; e += int64_t(gridDim.x) * kBS) {
    const uint32_t a = static_cast<uint32_t>(u[e]), b = static_cast<uint32_t>(v[e]);
    const uint32_t x = w ? static_cast<uint32_t>(int64_t(w[e]) - s.w_lo) : 0u;
    if (which == 1) {
      keys[e] = Pack(b, a, x, s);
    } else {
      keys[e] = Pack(a, b, x, s);
      if (which == 2) keys[m + e] = Pack(b, a, x, s);
    }
  }
}

// Sorted key i survives the squish: not a self-loop, and the first of its
// (row, destination) run.
__device__ __forceinline__ bool KeepAt(const uint64_t* __restrict__ keys, int64_t i, KeyShape s) {
  const uint64_t rd = keys[i] >> s.w_bits;
  const uint64_t mask = (uint64_t(1) << s.id_bits) - 1;
  if ((rd >> s.id_bits) == (rd & mask)) return false;
  return i == 0 || (keys[i - 1] >> s.w_bits) != rd;
}

__global__ __launch_bounds__(kBS) void BdMark(const uint64_t* __restrict__ keys, int64_t cnt, KeyShape s,
                                              uint32_t* __restrict__ flags) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < cnt; i += int64_t(gridDim.x) * kBS)
    flags[i] = KeepAt(keys, i, s) ? 1u : 0u;
}

// Writes the kept entries at their scanned slots (width 2: {dest, weight}
// pairs) and adds each row's size into deg[row]: its last sorted entry adds
// its end slot, its first subtracts its start slot.
__global__ __launch_bounds__(kBS) void BdEmit(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ pos,
                                              int64_t cnt, KeyShape s, int width, int32_t* __restrict__ dst,
                                              unsigned long long* __restrict__ deg) {
  const uint64_t dmask = (uint64_t(1) << s.id_bits) - 1;
  const uint64_t wmask = s.w_bits ? (uint64_t(1) << s.w_bits) - 1 : 0;
  const int row_shift = s.id_bits + s.w_bits;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < cnt; i += int64_t(gridDim.x) * kBS) {
    const uint64_t k = keys[i];
    const uint64_t row = k >> row_shift;
    const uint64_t p = pos[i];
    const bool keep = KeepAt(keys, i, s);
    if (keep) {
      const int32_t dest = static_cast<int32_t>((k >> s.w_bits) & dmask);
      if (width == 1) {
        dst[p] = dest;
      } else {
        dst[2 * p] = dest;
        dst[2 * p + 1] = static_cast<int32_t>(int64_t(k & wmask) + s.w_lo);
      }
    }
    if (i == 0 || (keys[i - 1] >> row_shift) != row) atomicAdd(&deg[row], 0ull - p);
    if (i + 1 == cnt || (keys[i + 1] >> row_shift) != row) atomicAdd(&deg[row], p + (keep ? 1 : 0));
  }
}

__global__ __launch_bounds__(kBS) void BdZero(unsigned long long* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) p[i] = 0;
}

__global__ __launch_bounds__(kBS) void BdMaxDegree(const int64_t* __restrict__ off, int64_t n,
                                                   unsigned long long* mx) {
  unsigned long long best = 0;
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    const unsigned long long d = static_cast<unsigned long long>(off[i + 1] - off[i]);
    best = d > best ? d : best;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(best, o);
    best = x > best ? x : best;
  }
  if (LaneId() == 0 && best) atomicMax(mx, best);
}

// Device memory freed on scope exit.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  explicit DevBuf(size_t b) : p(DeviceAlloc(b)), bytes(b) {}
  ~DevBuf() { release(); }
  void release() {
    DeviceFree(p, bytes);
    p = nullptr;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Temporaries of one side's build, allocated before the timed region.
struct SideBufs {
  DevBuf keys, tmp, flags, total;
  explicit SideBufs(int64_t cnt)
      : keys(sizeof(uint64_t) * (cnt > 0 ? cnt : 1)),
        tmp(sizeof(uint64_t) * (cnt > 0 ? cnt : 1)),
        flags(sizeof(uint32_t) * (cnt > 0 ? cnt : 1)),
        total(sizeof(uint32_t)) {}
};

// One side of the CSR (out-lists, in-lists, or both directions of a
// symmetrised graph) from the edge list on the device: fills off (n + 1)
// and dst (room for every unsquished entry) and returns the entries.
int64_t BuildSide(gg_graph* g, const int32_t* du, const int32_t* dv, const int32_t* dw, int64_t m, int which,
                  KeyShape s, int width, SideBufs& b, int64_t* off, int32_t* dst) {
  const int64_t n = g->n;
  const int64_t cnt = which == 2 ? 2 * m : m;
  const dim3 blk(kBS);
  const dim3 grid_m(GridFor(m, kBS, g->cu_count * 16)), grid_c(GridFor(cnt, kBS, g->cu_count * 16));
  if (m > 0)
    Run(g, "build_keys", (which == 2 ? 20.0 : 12.0) * m, BdKeys, grid_m, blk, du, dv, dw, m, s, which,
        b.keys.as<uint64_t>());
  uint64_t* sorted = RadixSort(g, "bd.sort", b.keys.as<uint64_t>(), b.tmp.as<uint64_t>(), cnt, 0,
                               2 * s.id_bits + s.w_bits);
  uint32_t entries32 = 0;
  GG_HIP(hipMemsetAsync(b.total.p, 0, sizeof(uint32_t), g->stream));
  unsigned long long* deg = reinterpret_cast<unsigned long long*>(off);
  Run(g, "build_zero", 8.0 * (n + 1), BdZero, dim3(GridFor(n + 1, kBS, g->cu_count * 4)), blk, deg, n + 1);
  if (cnt > 0) {
    Run(g, "build_mark", 12.0 * cnt, BdMark, grid_c, blk, static_cast<const uint64_t*>(sorted), cnt, s,
        b.flags.as<uint32_t>());
    DeviceScan<uint32_t>(g, "bd.flags", b.flags.as<uint32_t>(), cnt, b.total.as<uint32_t>());
    Run(g, "build_emit", 12.0 * cnt, BdEmit, grid_c, blk, static_cast<const uint64_t*>(sorted),
        static_cast<const uint32_t*>(b.flags.as<uint32_t>()), cnt, s, width, dst, deg);
  }
  DeviceScan<int64_t>(g, "bd.offsets", off, n + 1, static_cast<int64_t*>(nullptr));
  GG_HIP(hipMemcpyAsync(&entries32, b.total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  return entries32;
}

// ---- generation (src/generator.h) --------------------------------------
//
// Generator::MakeRMatEL, MakeUniformEL and InsertWeights (generator.h:
// 105-150, 170-187) as the host restatement (gap_gen.h) has them: edges in
// blocks of 2^18, each block's std::mt19937 seeded with seed + block index,
// so a block is one sequential stream -- here one wavefront (a 64-thread
// workgroup, so a barrier is a wavefront barrier).  The engine is the
// standard's mt19937 ([rand.predef]: the seeding recurrence, the 624-word
// twist, the tempering).  The twist runs in three phases, each parallel
// across the lanes: words [0, 227) read only old words, [227, 454) old words
// and the first phase's, [454, 624) the second phase's and word 0.  Each
// twist's 624 tempered outputs go to LDS, where the lanes turn them into
// edges or weights in stream order.
constexpr int kMtN = 624;
constexpr int64_t kGenEdges = int64_t(1) << 18;

__device__ __forceinline__ uint32_t MtMix(uint32_t xk, uint32_t xk1, uint32_t xm) {
  const uint32_t y = (xk & 0x80000000u) | (xk1 & 0x7fffffffu);
  return xm ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

__device__ __forceinline__ uint32_t MtTemper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// std::mt19937::seed(seed).
__device__ void MtSeed(uint32_t* mt, uint32_t seed) {
  if (threadIdx.x == 0) {
    uint32_t x = seed;
    mt[0] = x;
    for (int i = 1; i < kMtN; i++) {
      x = 1812433253u * (x ^ (x >> 30)) + static_cast<uint32_t>(i);
      mt[i] = x;
    }
  }
  __syncthreads();
}

template <int kLo, int kHi, int kRounds>
__device__ __forceinline__ void MtPhase(uint32_t* mt) {
  uint32_t t[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const int k = kLo + static_cast<int>(threadIdx.x) + 64 * r;
    if (k < kHi) {
      const int far = k < 227 ? k + 397 : k - 227;
      t[r] = MtMix(mt[k], mt[k + 1 == kMtN ? 0 : k + 1], mt[far]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const int k = kLo + static_cast<int>(threadIdx.x) + 64 * r;
    if (k < kHi) mt[k] = t[r];
  }
  __syncthreads();
}

// The next 624 outputs of the engine into out[].
__device__ void MtNext(uint32_t* mt, uint32_t* out) {
  MtPhase<0, 227, 4>(mt);
  MtPhase<227, 454, 4>(mt);
  MtPhase<454, kMtN, 3>(mt);
  for (int i = threadIdx.x; i < kMtN; i += 64) out[i] = MtTemper(mt[i]);
  __syncthreads();
}

// R-MAT edges (generator.h:122-150): `scale` draws per edge, most
// significant level first; endpoints through the permutation perm.  Edge e
// of a block ends in twist T when floor(T * 624 / scale) <= e <
// floor((T + 1) * 624 / scale); its draws lie in twists T - 1 and T, the
// two halves of out[].
__global__ __launch_bounds__(64) void GenRmat(int scale, int64_t m, int64_t seed, uint32_t a, uint32_t ab,
                                              uint32_t abc, const int32_t* __restrict__ perm,
                                              int32_t* __restrict__ u, int32_t* __restrict__ v) {
  __shared__ uint32_t mt[kMtN];
  __shared__ uint32_t out[2][kMtN];
  for (int64_t blk = blockIdx.x; blk * kGenEdges < m; blk += gridDim.x) {
    const int64_t lo = blk * kGenEdges;
    const int64_t be = m - lo < kGenEdges ? m - lo : kGenEdges;
    MtSeed(mt, static_cast<uint32_t>(seed + blk));
    int64_t done = 0;
    for (int64_t t = 0; done < be; t++) {
      MtNext(mt, out[t & 1]);
      int64_t hi = (t + 1) * kMtN / scale;
      if (hi > be) hi = be;
      for (int64_t e = done + threadIdx.x; e < hi; e += 64) {
        uint32_t s = 0, d = 0;
        for (int l = 0; l < scale; l++) {
          const int64_t p = e * scale + l;
          const int64_t tw = p / kMtN;
          const uint32_t r = out[tw & 1][p - tw * kMtN];
          s <<= 1;
          d <<= 1;
          if (r < ab) {
            d += r > a;
          } else {
            s += 1;
            d += r > abc;
          }
        }
        u[lo + e] = perm[s];
        v[lo + e] = perm[d];
      }
      done = hi;
      __syncthreads();  // twist t + 1 overwrites twist t - 1's half
    }
  }
}

// Uniform edges (generator.h:105-120, no permutation): two draws per edge,
// the destination first (the order g++ evaluates Edge(udist(), udist())
// in); for n = 2^scale the bounded draw is the low bits, never a rejection.
__global__ __launch_bounds__(64) void GenUniform(int scale, int64_t m, int64_t seed, int32_t* __restrict__ u,
                                                 int32_t* __restrict__ v) {
  __shared__ uint32_t mt[kMtN];
  __shared__ uint32_t out[kMtN];
  const uint32_t mask = (uint32_t(1) << scale) - 1u;
  for (int64_t blk = blockIdx.x; blk * kGenEdges < m; blk += gridDim.x) {
    const int64_t lo = blk * kGenEdges;
    const int64_t be = m - lo < kGenEdges ? m - lo : kGenEdges;
    MtSeed(mt, static_cast<uint32_t>(seed + blk));
    for (int64_t first = 0; first < be; first += kMtN / 2) {
      MtNext(mt, out);
      for (int j = threadIdx.x; j < kMtN / 2 && first + j < be; j += 64) {
        v[lo + first + j] = static_cast<int32_t>(out[2 * j] & mask);
        u[lo + first + j] = static_cast<int32_t>(out[2 * j + 1] & mask);
      }
      __syncthreads();
    }
  }
}

// Weights uniform in [1, max_weight] (generator.h:170-187): one bounded
// draw per weight, rejecting draws at or above the cutoff (for 255:
// 0xFFFFFFFF only), so a weight's index is the count of accepted draws
// before it in the block's stream (a ballot per 64 outputs).
__global__ __launch_bounds__(64) void GenWeights(int64_t m, int64_t seed, uint32_t mod, uint32_t cutoff,
                                                 int32_t* __restrict__ w) {
  __shared__ uint32_t mt[kMtN];
  __shared__ uint32_t out[kMtN];
  const int lane = static_cast<int>(threadIdx.x);
  for (int64_t blk = blockIdx.x; blk * kGenEdges < m; blk += gridDim.x) {
    const int64_t lo = blk * kGenEdges;
    const int64_t be = m - lo < kGenEdges ? m - lo : kGenEdges;
    MtSeed(mt, static_cast<uint32_t>(seed + blk));
    int64_t base = 0;
    while (base < be) {
      MtNext(mt, out);
      for (int q = 0; q * 64 < kMtN; q++) {
        const int i = q * 64 + lane;
        const uint32_t r = i < kMtN ? out[i] : 0u;
        const bool ok = i < kMtN && (cutoff == 0 || r < cutoff);
        const uint64_t mk = __ballot(ok);
        const int64_t idx = base + __popcll(mk & LanesBelow(~0ull));
        if (ok && idx < be) w[lo + idx] = static_cast<int32_t>(r % mod) + 1;
        base += __popcll(mk);
      }
      __syncthreads();
    }
  }
}

// The squished CSR of the edge list du/dv/dw (m edges, device memory) in a
// new handle on `device`; *num_nodes gets n.  stats->device_ms covers the
// build from the edge list in HBM.
gg_graph* BuildFromDevice(int device, const int32_t* du, const int32_t* dv, const int32_t* dw, int64_t m,
                          bool symmetrize, bool weighted, gg_stats* stats, const WallTimer& wt) {
  const bool use_w = weighted && dw != nullptr;
  const int width = weighted ? 2 : 1;
  Range hr{INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
  if (m > 0) {
    DevBuf dr(sizeof(Range));
    GG_HIP(hipMemcpy(dr.p, &hr, sizeof(Range), hipMemcpyHostToDevice));
    BdRange<<<GridFor(m, kBS, 4096), kBS>>>(du, dv, use_w ? dw : nullptr, m, dr.as<Range>());
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpy(&hr, dr.p, sizeof(Range), hipMemcpyDeviceToHost));
    GG_REQUIRE(hr.lo_id >= 0, GG_EINVAL, "negative vertex id in the edge list");
  }
  const int64_t n = m > 0 ? int64_t(hr.hi_id) + 1 : 0;
  KeyShape s;
  s.id_bits = BitsFor(n > 0 ? static_cast<uint64_t>(n - 1) : 0);
  s.w_lo = use_w ? hr.lo_w : 1;
  s.w_bits = use_w ? BitsFor(static_cast<uint64_t>(int64_t(hr.hi_w) - hr.lo_w)) : 0;
  GG_REQUIRE(2 * s.id_bits + s.w_bits <= 64, GG_ENOTSUP,
             "vertex ids and weight range too wide to sort as one 64-bit key");

  // Everything is allocated before the timed region: the lists with room
  // for every unsquished entry (the squish drops only duplicates and
  // self-loops), the keys, flags and scan space.
  const int64_t cnt = symmetrize ? 2 * m : m;
  gg_graph* g = NewDeviceGraph(device, n, 0);
  try {
    g->weighted = weighted;
    DeviceFree(g->out_dst, 0);
    g->out_dst = g->in_dst = nullptr;
    g->out_dst = static_cast<int32_t*>(DeviceAlloc(sizeof(int32_t) * width * cnt));
    if (!symmetrize) {
      // Directed from here on, so that a failure frees the in-lists too.
      g->in_off = static_cast<int64_t*>(DeviceAlloc(sizeof(int64_t) * (n + 1)));
      g->directed = true;
      g->in_dst = static_cast<int32_t*>(DeviceAlloc(sizeof(int32_t) * width * cnt));
    }
    SideBufs bufs(cnt);
    const int64_t tiles = (cnt + kRsTile - 1) / kRsTile;
    (void)g->Scratch<uint32_t>("bd.sort.hist", tiles * 256);
    BeginTiming(g);
    const int32_t* pw = use_w ? dw : nullptr;
    g->out_entries = BuildSide(g, du, dv, pw, m, symmetrize ? 2 : 0, s, width, bufs, g->out_off, g->out_dst);
    g->graph_bytes = sizeof(int64_t) * (n + 1) + sizeof(int32_t) * width * g->out_entries;
    if (!symmetrize) {
      g->in_entries = BuildSide(g, du, dv, pw, m, 1, s, width, bufs, g->in_off, g->in_dst);
      g->graph_bytes += sizeof(int64_t) * (n + 1) + sizeof(int32_t) * width * g->in_entries;
    } else {
      g->in_off = g->out_off;
      g->in_dst = g->out_dst;
      g->in_entries = g->out_entries;
    }
    DevBuf mx(sizeof(unsigned long long));
    unsigned long long hmx = 0;
    GG_HIP(hipMemsetAsync(mx.p, 0, sizeof(hmx), g->stream));
    Run(g, "build_max_degree", 8.0 * (n + 1), BdMaxDegree, dim3(GridFor(n, kBS, g->cu_count * 4)), dim3(kBS),
        static_cast<const int64_t*>(g->out_off), n, mx.as<unsigned long long>());
    if (g->directed)
      Run(g, "build_max_degree", 8.0 * (n + 1), BdMaxDegree, dim3(GridFor(n, kBS, g->cu_count * 4)), dim3(kBS),
          static_cast<const int64_t*>(g->in_off), n, mx.as<unsigned long long>());
    GG_HIP(hipMemcpyAsync(&hmx, mx.p, sizeof(hmx), hipMemcpyDeviceToHost, g->stream));
    FinishTiming(g, stats, wt);
    g->max_degree = static_cast<int64_t>(hmx);
    // Outside the timed region: each list array down to its exact size.
    auto shrink = [&](int32_t** p, int64_t entries) {
      if (entries == cnt) return;
      int32_t* q = static_cast<int32_t*>(DeviceAlloc(sizeof(int32_t) * width * entries));
      if (entries) GG_HIP(hipMemcpy(q, *p, sizeof(int32_t) * width * entries, hipMemcpyDeviceToDevice));
      DeviceFree(*p, sizeof(int32_t) * width * cnt);
      *p = q;
    };
    shrink(&g->out_dst, g->out_entries);
    if (g->directed)
      shrink(&g->in_dst, g->in_entries);
    else
      g->in_dst = g->out_dst;
    g->ws.Release();  // the sort's and scans' workspaces
  } catch (...) {
    gg_graph_free(g);
    throw;
  }
  if (stats) stats->reached_edges = g->out_entries;
  return g;
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_graph_build(const int32_t* u, const int32_t* v, const int32_t* w, int64_t m, int symmetrize,
                              int weighted, gg_graph** out, int64_t* num_nodes, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(out != nullptr, GG_EINVAL, "null output handle");
    *out = nullptr;
    GG_REQUIRE(GpuInitialised(), GG_ENODEV, "gg_init has not been called");
    GG_REQUIRE(m >= 0, GG_EINVAL, "negative edge count");
    GG_REQUIRE(m == 0 || (u && v), GG_EINVAL, "null edge arrays");
    GG_REQUIRE((symmetrize ? 2 * m : m) < (int64_t(1) << 32), GG_ENOTSUP, "2^32 or more list entries");
    int device = 0;
    GG_HIP(hipGetDevice(&device));
    const bool use_w = weighted && w != nullptr;
    DevBuf du(sizeof(int32_t) * (m > 0 ? m : 1)), dv(sizeof(int32_t) * (m > 0 ? m : 1));
    DevBuf dw(use_w ? sizeof(int32_t) * (m > 0 ? m : 1) : 0);
    if (m > 0) {
      GG_HIP(hipMemcpy(du.p, u, sizeof(int32_t) * m, hipMemcpyHostToDevice));
      GG_HIP(hipMemcpy(dv.p, v, sizeof(int32_t) * m, hipMemcpyHostToDevice));
      if (use_w) GG_HIP(hipMemcpy(dw.p, w, sizeof(int32_t) * m, hipMemcpyHostToDevice));
    }
    gg_graph* g = BuildFromDevice(device, du.as<int32_t>(), dv.as<int32_t>(), use_w ? dw.as<int32_t>() : nullptr, m,
                                  symmetrize != 0, weighted != 0, stats, wt);
    *out = g;
    if (num_nodes) *num_nodes = g->n;
    if (stats) stats->total_ms = wt.ms();
  });
}

extern "C" int gg_graph_generate(int kind, int scale, int degree, int64_t seed, int weighted,
                                 const int32_t* perm, gg_graph** out, int64_t* num_nodes, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(out != nullptr, GG_EINVAL, "null output handle");
    *out = nullptr;
    GG_REQUIRE(GpuInitialised(), GG_ENODEV, "gg_init has not been called");
    GG_REQUIRE(kind == 0 || kind == 1, GG_ENOTSUP, "device generation covers Kronecker (0) and uniform (1)");
    GG_REQUIRE(scale >= 1 && scale <= 30 && degree >= 1, GG_EINVAL, "scale in [1, 30] and degree >= 1");
    GG_REQUIRE(kind == 1 || perm != nullptr, GG_EINVAL, "Kronecker graphs need the vertex permutation");
    const int64_t n = int64_t(1) << scale, m = n * degree;
    GG_REQUIRE(2 * m < (int64_t(1) << 32), GG_ENOTSUP, "2^32 or more list entries");
    int device = 0;
    GG_HIP(hipGetDevice(&device));
    DevBuf du(sizeof(int32_t) * m), dv(sizeof(int32_t) * m), dw(weighted ? sizeof(int32_t) * m : 0);
    DevBuf dp(kind == 0 ? sizeof(int32_t) * n : 0);
    if (kind == 0) GG_HIP(hipMemcpy(dp.p, perm, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float gen_ms = 0;
    auto cleanup = [&] {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (st) (void)hipStreamDestroy(st);
    };
    try {
      GG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      GG_HIP(hipEventCreate(&e0));
      GG_HIP(hipEventCreate(&e1));
      const unsigned grid = GridFor((m + kGenEdges - 1) / kGenEdges, 1, 65536);
      GG_HIP(hipEventRecord(e0, st));
      if (kind == 0) {
        // The thresholds exactly as the host computes them (generator.h:124).
        const uint32_t kMax = 0xFFFFFFFFu;
        const uint32_t a = static_cast<uint32_t>(0.57 * kMax), b = static_cast<uint32_t>(0.19 * kMax),
                       c = static_cast<uint32_t>(0.19 * kMax);
        GenRmat<<<grid, 64, 0, st>>>(scale, m, seed, a, a + b, a + b + c, dp.as<int32_t>(), du.as<int32_t>(),
                                     dv.as<int32_t>());
      } else {
        GenUniform<<<grid, 64, 0, st>>>(scale, m, seed, du.as<int32_t>(), dv.as<int32_t>());
      }
      if (weighted) {
        // BoundedDraw over [0, 254] (gap_gen.h): mod 255, cutoff 2^32 - 1 - (2^32 - 1) % 255.
        const uint32_t mod = 255, rem = 0xFFFFFFFFu % mod;
        const uint32_t cutoff = rem == mod - 1 ? 0u : 0xFFFFFFFFu - rem;
        GenWeights<<<grid, 64, 0, st>>>(m, seed, mod, cutoff, dw.as<int32_t>());
      }
      GG_HIP(hipGetLastError());
      GG_HIP(hipEventRecord(e1, st));
      GG_HIP(hipEventSynchronize(e1));
      GG_HIP(hipEventElapsedTime(&gen_ms, e0, e1));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
    dp.release();
    gg_graph* g = BuildFromDevice(device, du.as<int32_t>(), dv.as<int32_t>(), weighted ? dw.as<int32_t>() : nullptr,
                                  m, true, weighted != 0, stats, wt);
    *out = g;
    if (num_nodes) *num_nodes = g->n;
    if (stats) {
      stats->device_ms += gen_ms;
      stats->total_ms = wt.ms();
    }
  });
}

extern "C" int gg_graph_copy_in(const gg_graph* g, int64_t* offsets, int32_t* dests) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_HIP(hipSetDevice(g->device));
    GG_HIP(hipStreamSynchronize(g->stream));
    const int width = g->weighted ? 2 : 1;
    if (offsets) GG_HIP(hipMemcpy(offsets, g->in_off, sizeof(int64_t) * (g->n + 1), hipMemcpyDeviceToHost));
    if (dests && g->in_entries)
      GG_HIP(hipMemcpy(dests, g->in_dst, sizeof(int32_t) * width * g->in_entries, hipMemcpyDeviceToHost));
  });
}
