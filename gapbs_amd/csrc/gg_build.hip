// gg_graph_build: the squished CSR of an edge list, built on the device.
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
// are freed before the call returns.
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

// One side of the CSR (out-lists, in-lists, or both directions of a
// symmetrised graph) from the edge list on the device: fills *off (n + 1,
// allocated by the caller) and returns the entries and their array.
int64_t BuildSide(gg_graph* g, const int32_t* du, const int32_t* dv, const int32_t* dw, int64_t m, int which,
                  KeyShape s, int width, int64_t* off, int32_t** dst) {
  const int64_t n = g->n;
  const int64_t cnt = which == 2 ? 2 * m : m;
  const dim3 blk(kBS);
  DevBuf keys(sizeof(uint64_t) * (cnt > 0 ? cnt : 1)), tmp(sizeof(uint64_t) * (cnt > 0 ? cnt : 1));
  DevBuf flags(sizeof(uint32_t) * (cnt > 0 ? cnt : 1)), total(sizeof(uint32_t));
  const dim3 grid_m(GridFor(m, kBS, g->cu_count * 16)), grid_c(GridFor(cnt, kBS, g->cu_count * 16));
  if (m > 0)
    Run(g, "build_keys", (which == 2 ? 20.0 : 12.0) * m, BdKeys, grid_m, blk, du, dv, dw, m, s, which,
        keys.as<uint64_t>());
  uint64_t* sorted = RadixSort(g, "bd.sort", keys.as<uint64_t>(), tmp.as<uint64_t>(), cnt, 0,
                               2 * s.id_bits + s.w_bits);
  uint32_t entries32 = 0;
  GG_HIP(hipMemsetAsync(total.p, 0, sizeof(uint32_t), g->stream));
  if (cnt > 0) {
    Run(g, "build_mark", 12.0 * cnt, BdMark, grid_c, blk, static_cast<const uint64_t*>(sorted), cnt, s,
        flags.as<uint32_t>());
    DeviceScan<uint32_t>(g, "bd.flags", flags.as<uint32_t>(), cnt, total.as<uint32_t>());
  }
  GG_HIP(hipMemcpyAsync(&entries32, total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  const int64_t entries = entries32;
  *dst = static_cast<int32_t*>(DeviceAlloc(sizeof(int32_t) * width * entries));
  unsigned long long* deg = reinterpret_cast<unsigned long long*>(off);
  Run(g, "build_zero", 8.0 * (n + 1), BdZero, dim3(GridFor(n + 1, kBS, g->cu_count * 4)), blk, deg, n + 1);
  if (cnt > 0)
    Run(g, "build_emit", 12.0 * cnt + 4.0 * width * entries, BdEmit, grid_c, blk,
        static_cast<const uint64_t*>(sorted), static_cast<const uint32_t*>(flags.as<uint32_t>()), cnt, s, width,
        *dst, deg);
  DeviceScan<int64_t>(g, "bd.offsets", off, n + 1, static_cast<int64_t*>(nullptr));
  return entries;
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
    const int64_t cnt = symmetrize ? 2 * m : m;
    GG_REQUIRE(cnt < (int64_t(1) << 32), GG_ENOTSUP, "2^32 or more list entries");
    int device = 0;
    GG_HIP(hipGetDevice(&device));
    const bool use_w = weighted && w != nullptr;
    const int width = weighted ? 2 : 1;

    // The edge list in HBM and its ranges.
    DevBuf du(sizeof(int32_t) * (m > 0 ? m : 1)), dv(sizeof(int32_t) * (m > 0 ? m : 1));
    DevBuf dw(use_w ? sizeof(int32_t) * (m > 0 ? m : 1) : 0);
    if (m > 0) {
      GG_HIP(hipMemcpy(du.p, u, sizeof(int32_t) * m, hipMemcpyHostToDevice));
      GG_HIP(hipMemcpy(dv.p, v, sizeof(int32_t) * m, hipMemcpyHostToDevice));
      if (use_w) GG_HIP(hipMemcpy(dw.p, w, sizeof(int32_t) * m, hipMemcpyHostToDevice));
    }
    Range hr{INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    if (m > 0) {
      DevBuf dr(sizeof(Range));
      GG_HIP(hipMemcpy(dr.p, &hr, sizeof(Range), hipMemcpyHostToDevice));
      BdRange<<<GridFor(m, kBS, 4096), kBS>>>(du.as<int32_t>(), dv.as<int32_t>(),
                                              use_w ? dw.as<int32_t>() : nullptr, m, dr.as<Range>());
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

    gg_graph* g = NewDeviceGraph(device, n, 0);
    try {
      g->weighted = weighted != 0;
      DeviceFree(g->out_dst, 0);
      g->out_dst = g->in_dst = nullptr;
      BeginTiming(g);
      const int32_t* pw = use_w ? dw.as<int32_t>() : nullptr;
      g->out_entries = BuildSide(g, du.as<int32_t>(), dv.as<int32_t>(), pw, m, symmetrize ? 2 : 0, s, width,
                                 g->out_off, &g->out_dst);
      g->graph_bytes = sizeof(int64_t) * (n + 1) + sizeof(int32_t) * width * g->out_entries;
      if (!symmetrize) {
        // Directed from here on, so that a failure frees the in-lists too.
        g->in_off = static_cast<int64_t*>(DeviceAlloc(sizeof(int64_t) * (n + 1)));
        g->in_dst = nullptr;
        g->directed = true;
        g->in_entries = BuildSide(g, du.as<int32_t>(), dv.as<int32_t>(), pw, m, 1, s, width, g->in_off, &g->in_dst);
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
      g->ws.Release();  // the sort's and scans' workspaces
    } catch (...) {
      gg_graph_free(g);
      throw;
    }
    *out = g;
    if (num_nodes) *num_nodes = n;
    if (stats) {
      stats->total_ms = wt.ms();
      stats->reached_edges = g->out_entries;
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
