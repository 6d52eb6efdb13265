// Vertex ids across the C-ABI for degree-ordered graphs
// (gg_graph_order_by_degree).  Callers always name vertices by their
// uploaded ids and get per-vertex results indexed by them; a degree-ordered
// graph keeps its lists in device ids, so roots are mapped on the way in
// (new_id) and results on the way out (a gather through new_id, plus old_id
// for values that are vertex ids).  On a graph in uploaded order every
// helper is a plain copy.
#include <type_traits>

#include "gg_common.h"

namespace gg {
namespace {

constexpr int kBS = 256;

__global__ __launch_bounds__(kBS) void OrMapIds(int32_t* __restrict__ d, int64_t k,
                                                const int32_t* __restrict__ new_id) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < k; i += int64_t(gridDim.x) * kBS)
    d[i] = new_id[d[i]];
}

// out[v] = src[new_id[v]] for original ids v.
template <class T>
__global__ __launch_bounds__(kBS) void OrGather(const T* __restrict__ src, const int32_t* __restrict__ new_id,
                                                int64_t n, T* __restrict__ out) {
  for (int64_t v = blockIdx.x * int64_t(kBS) + threadIdx.x; v < n; v += int64_t(gridDim.x) * kBS)
    out[v] = src[new_id[v]];
}

// The same for values that are device ids (mapped back; negative values
// are kept).
__global__ __launch_bounds__(kBS) void OrGatherIds(const int32_t* __restrict__ src,
                                                   const int32_t* __restrict__ new_id,
                                                   const int32_t* __restrict__ old_id, int64_t n,
                                                   int32_t* __restrict__ out) {
  for (int64_t v = blockIdx.x * int64_t(kBS) + threadIdx.x; v < n; v += int64_t(gridDim.x) * kBS) {
    const int32_t x = src[new_id[v]];
    out[v] = x >= 0 ? old_id[x] : x;
  }
}

__global__ __launch_bounds__(kBS) void OrFill(int32_t* __restrict__ p, int64_t n, int32_t value) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) p[i] = value;
}

// least[c] = min original id over the vertices labelled c.  The lanes of a
// wavefront that share lane 0's label (most of them, for a giant
// component) reduce first and post one atomic; an atomic is skipped when
// the stored minimum is already no larger.
__global__ __launch_bounds__(kBS) void OrComponentMin(const int32_t* __restrict__ comp,
                                                      const int32_t* __restrict__ old_id, int64_t n,
                                                      int32_t* least) {
  const int lane = LaneId();
  const int64_t stride = int64_t(gridDim.x) * kBS;
  // Every lane of a wavefront runs the same number of iterations (the bound
  // is rounded up to whole wavefronts) so the shuffles below see all lanes.
  const int64_t bound = (n + kWave - 1) / kWave * kWave;
  for (int64_t p = blockIdx.x * int64_t(kBS) + threadIdx.x; p < bound; p += stride) {
    const bool valid = p < n;
    const int32_t c = valid ? comp[p] : -1;
    const int32_t o = valid ? old_id[p] : INT32_MAX;
    const int32_t c0 = __shfl(c, 0);
    const bool same = valid && c == c0;
    int32_t m = same ? o : INT32_MAX;
    for (int off = 32; off > 0; off >>= 1) {
      const int32_t t = __shfl_xor(m, off);
      m = t < m ? t : m;
    }
    if (lane == 0 && c0 >= 0 && m < least[c0]) atomicMin(&least[c0], m);
    if (valid && !same && o < least[c]) atomicMin(&least[c], o);
  }
}

__global__ __launch_bounds__(kBS) void OrComponentOut(const int32_t* __restrict__ comp,
                                                      const int32_t* __restrict__ least,
                                                      const int32_t* __restrict__ new_id, int64_t n,
                                                      int32_t* __restrict__ out) {
  for (int64_t v = blockIdx.x * int64_t(kBS) + threadIdx.x; v < n; v += int64_t(gridDim.x) * kBS)
    out[v] = least[comp[new_id[v]]];
}

template <class T>
void CopyOutVertex(gg_graph* g, const T* src, T* host, bool ids) {
  const int64_t n = g->n;
  if (!g->new_id) {
    GG_HIP(hipMemcpyAsync(host, src, sizeof(T) * n, hipMemcpyDeviceToHost, g->stream));
  } else {
    T* tmp = g->Scratch<T>("order.out", n);
    const dim3 grid(GridFor(n, kBS, g->cu_count * 16)), blk(kBS);
    if constexpr (std::is_same<T, int32_t>::value) {
      if (ids) {
        hipLaunchKernelGGL(OrGatherIds, grid, blk, 0, g->stream, src, static_cast<const int32_t*>(g->new_id),
                           static_cast<const int32_t*>(g->old_id), n, tmp);
      } else {
        hipLaunchKernelGGL(OrGather<T>, grid, blk, 0, g->stream, src, static_cast<const int32_t*>(g->new_id), n,
                           tmp);
      }
    } else {
      hipLaunchKernelGGL(OrGather<T>, grid, blk, 0, g->stream, src, static_cast<const int32_t*>(g->new_id), n, tmp);
    }
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpyAsync(host, tmp, sizeof(T) * n, hipMemcpyDeviceToHost, g->stream));
  }
  GG_HIP(hipStreamSynchronize(g->stream));
}

}  // namespace

int32_t DeviceVertex(gg_graph* g, int32_t v) {
  if (!g->new_id) return v;
  int32_t d = 0;
  GG_HIP(hipMemcpyAsync(&d, g->new_id + v, sizeof(d), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  return d;
}

void DeviceVertices(gg_graph* g, int32_t* d, int64_t k) {
  if (!g->new_id || k <= 0) return;
  hipLaunchKernelGGL(OrMapIds, dim3(GridFor(k, kBS, g->cu_count * 4)), dim3(kBS), 0, g->stream, d, k,
                     static_cast<const int32_t*>(g->new_id));
  GG_HIP(hipGetLastError());
}

void CopyOutVertexI32(gg_graph* g, const int32_t* src, int32_t* host, bool ids) {
  CopyOutVertex<int32_t>(g, src, host, ids);
}

void CopyOutVertexF32(gg_graph* g, const float* src, float* host) { CopyOutVertex<float>(g, src, host, false); }

void CopyOutComponentsWith(gg_graph* g, const int32_t* comp, int32_t* host, int32_t* least, int32_t* tmp) {
  const int64_t n = g->n;
  if (!g->new_id) {
    GG_HIP(hipMemcpyAsync(host, comp, sizeof(int32_t) * n, hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
    return;
  }
  const dim3 grid(GridFor(n, kBS, g->cu_count * 16)), blk(kBS);
  hipLaunchKernelGGL(OrFill, grid, blk, 0, g->stream, least, n, INT32_MAX);
  hipLaunchKernelGGL(OrComponentMin, grid, blk, 0, g->stream, comp, static_cast<const int32_t*>(g->old_id), n,
                     least);
  hipLaunchKernelGGL(OrComponentOut, grid, blk, 0, g->stream, comp, static_cast<const int32_t*>(least),
                     static_cast<const int32_t*>(g->new_id), n, tmp);
  GG_HIP(hipGetLastError());
  GG_HIP(hipMemcpyAsync(host, tmp, sizeof(int32_t) * n, hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
}

void CopyOutComponents(gg_graph* g, const int32_t* comp, int32_t* host) {
  if (!g->new_id) {
    CopyOutVertex<int32_t>(g, comp, host, false);
    return;
  }
  CopyOutComponentsWith(g, comp, host, g->Scratch<int32_t>("order.least", g->n),
                        g->Scratch<int32_t>("order.out", g->n));
}

}  // namespace gg

using namespace gg;

extern "C" int gg_graph_vertex_order(gg_graph* g, int32_t* new_id) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(new_id != nullptr, GG_EINVAL, "null output");
    GG_HIP(hipSetDevice(g->device));
    if (!g->new_id) {
      for (int64_t v = 0; v < g->n; v++) new_id[v] = static_cast<int32_t>(v);
      return;
    }
    GG_HIP(hipMemcpyAsync(new_id, g->new_id, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, g->stream));
    GG_HIP(hipStreamSynchronize(g->stream));
  });
}
