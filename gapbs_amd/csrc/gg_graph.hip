// Device initialisation, graph upload and the shared runtime pieces of the
// kernel library.  The graph is copied into HBM once per handle, outside any
// timed region, like the reference's graph build (src/benchmark.h:107-109
// times only the kernel).
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <mutex>

#include "gg_common.h"

namespace gg {

namespace {
thread_local std::string t_error;
std::mutex g_alloc_mu;
size_t g_allocated = 0;
bool g_initialised = false;

size_t MemCap() {
  const char* s = std::getenv("GAPGPU_MEM_CAP_BYTES");
  if (!s || !*s) return 0;
  return static_cast<size_t>(std::strtoull(s, nullptr, 10));
}
}  // namespace

void SetError(const std::string& msg) { t_error = msg; }

void* DeviceAlloc(size_t bytes) {
  if (bytes == 0) bytes = 256;
  {
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    const size_t cap = MemCap();
    if (cap && g_allocated + bytes > cap)
      throw Failure(GG_ENOMEM, "device allocation of " + std::to_string(bytes) +
                                   " bytes exceeds GAPGPU_MEM_CAP_BYTES");
  }
  void* p = nullptr;
  GG_HIP(hipMalloc(&p, bytes));
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocated += bytes;
  return p;
}

void DeviceFree(void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  if (bytes == 0) bytes = 256;
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocated -= bytes < g_allocated ? bytes : g_allocated;
}

void* Workspace::Get(const std::string& name, size_t bytes) {
  Slot& s = slots[name];
  if (s.p && s.bytes >= bytes) return s.p;
  DeviceFree(s.p, s.bytes);
  s.p = nullptr;
  s.bytes = 0;
  // Round up so that small growth does not reallocate every call.
  const size_t rounded = (bytes + (1u << 20) - 1) & ~size_t((1u << 20) - 1);
  s.p = DeviceAlloc(rounded);
  s.bytes = rounded;
  return s.p;
}

void Workspace::Release() {
  for (auto& kv : slots) DeviceFree(kv.second.p, kv.second.bytes);
  slots.clear();
}

void Workspace::Drop(const std::string& prefix) {
  for (auto it = slots.begin(); it != slots.end();) {
    if (it->first.compare(0, prefix.size(), prefix) == 0) {
      DeviceFree(it->second.p, it->second.bytes);
      it = slots.erase(it);
    } else {
      ++it;
    }
  }
}

size_t Workspace::Bytes() const {
  size_t b = 0;
  for (const auto& kv : slots) b += kv.second.bytes;
  return b;
}

void MailInit(gg_graph* g) {
  if (g->mail_host) return;
  void* p = nullptr;
  GG_HIP(hipHostMalloc(&p, sizeof(unsigned long long) * (kMailWords + 1),
                       hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(p, 0, sizeof(unsigned long long) * (kMailWords + 1));
  void* d = nullptr;
  GG_HIP(hipHostGetDevicePointer(&d, p, 0));
  g->mail_host = static_cast<volatile unsigned long long*>(p);
  g->mail_dev = static_cast<unsigned long long*>(d);
}

void MailWait(gg_graph* g, unsigned long long seq, unsigned long long* values, int k) {
  auto seen = [&] { return __atomic_load_n(&g->mail_host[0], __ATOMIC_ACQUIRE) >= seq; };
  const auto t0 = std::chrono::steady_clock::now();
  while (!seen()) {
    __builtin_ia32_pause();
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      // Long step (or a failure): block on the stream instead of spinning.
      GG_HIP(hipStreamSynchronize(g->stream));
      GG_REQUIRE(seen(), GG_EHIP, "device did not post its step results");
      break;
    }
  }
  for (int i = 0; i < k; i++) values[i] = g->mail_host[1 + i];
}

void BeginTiming(gg_graph* g) { GG_HIP(hipEventRecord(g->ev_start, g->stream)); }

void EnsureSideStream(gg_graph* g) {
  if (g->side) return;
  GG_HIP(hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking));
  GG_HIP(hipStreamCreateWithFlags(&g->side2, hipStreamNonBlocking));
  GG_HIP(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
  GG_HIP(hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming));
  GG_HIP(hipEventCreateWithFlags(&g->ev_join2, hipEventDisableTiming));
}

void FinishTiming(gg_graph* g, gg_stats* s, const WallTimer& wt) {
  GG_HIP(hipEventRecord(g->ev_stop, g->stream));
  GG_HIP(hipEventSynchronize(g->ev_stop));
  SettleProfile(g);
  if (s) {
    float ms = 0;
    GG_HIP(hipEventElapsedTime(&ms, g->ev_start, g->ev_stop));
    s->device_ms = ms;
    s->total_ms = wt.ms();
  }
}

int ProfileSlot(gg_graph* g, const char* name, double bytes, hipEvent_t* a, hipEvent_t* b) {
  hipEvent_t ev[2];
  for (hipEvent_t& e : ev) {
    if (g->event_pool.empty()) {
      GG_HIP(hipEventCreate(&e));
    } else {
      e = g->event_pool.back();
      g->event_pool.pop_back();
    }
  }
  g->pending.push_back(gg_graph::PendingLaunch{name, ev[0], ev[1], bytes});
  *a = ev[0];
  *b = ev[1];
  return static_cast<int>(g->pending.size()) - 1;
}

std::vector<unsigned long long> ReadWork(gg_graph* g, const unsigned long long* work, int kinds) {
  std::vector<unsigned long long> h(static_cast<size_t>(kinds) * kWorkCounters);
  GG_HIP(hipMemcpyAsync(h.data(), work, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, g->stream));
  GG_HIP(hipStreamSynchronize(g->stream));
  if (g->side) GG_HIP(hipStreamSynchronize(g->side));
  if (g->side2) GG_HIP(hipStreamSynchronize(g->side2));
  std::vector<unsigned long long> out(2 * kinds, 0);
  for (int k = 0; k < kinds; k++)
    for (int s = 0; s < kWorkStripes; s++) {
      out[2 * k] += h[k * kWorkCounters + s];
      out[2 * k + 1] += h[k * kWorkCounters + kWorkStripes + s];
    }
  return out;
}

void CreditBytes(gg_graph* g, int idx, double bytes) {
  if (idx >= 0) g->pending[idx].bytes += bytes;
}

void SettleProfile(gg_graph* g) {
  for (auto& p : g->pending) {
    float ms = 0;
    GG_HIP(hipEventSynchronize(p.b));
    GG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& t = g->kernel_totals[p.name];
    t.launches++;
    t.ms += ms;
    t.bytes += p.bytes;
    g->event_pool.push_back(p.a);
    g->event_pool.push_back(p.b);
  }
  g->pending.clear();
}

namespace {

// Upload-time validation: offsets non-decreasing, every destination in
// [0, n).  Kernels index per-vertex arrays with destinations unchecked, so a
// malformed graph is rejected here instead of faulting later.
__global__ __launch_bounds__(256) void CheckOffsets(const int64_t* __restrict__ off, int64_t n,
                                                    unsigned long long* max_deg, int* bad) {
  __shared__ unsigned long long s_max[4];
  unsigned long long mx = 0;
  bool neg = false;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t d = off[i + 1] - off[i];
    neg |= d < 0;
    if (d > 0 && static_cast<unsigned long long>(d) > mx) mx = static_cast<unsigned long long>(d);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(mx, o);
    mx = t > mx ? t : mx;
  }
  if (LaneId() == 0) s_max[threadIdx.x / 64] = mx;
  if (__ballot(neg) && LaneId() == 0) *bad = 1;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) mx = s_max[w] > mx ? s_max[w] : mx;
    mx = s_max[0] > mx ? s_max[0] : mx;
    if (mx) atomicMax(max_deg, mx);
  }
}

__global__ __launch_bounds__(256) void CheckDests(const int32_t* __restrict__ dst, int64_t entries,
                                                  int width, int64_t n, int* bad) {
  bool out = false;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < entries;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t v = dst[i * width];
    out |= v < 0 || v >= n;
  }
  if (__ballot(out) && LaneId() == 0) *bad = 1;
}

void ValidateSide(const int64_t* off, const int32_t* dst, int64_t entries, int width, int64_t n,
                  int64_t* max_degree) {
  struct Flags {
    unsigned long long max_deg;
    int bad;
  };
  Flags* d = static_cast<Flags*>(DeviceAlloc(sizeof(Flags)));
  Flags h{0, 0};
  try {
    GG_HIP(hipMemcpy(d, &h, sizeof(Flags), hipMemcpyHostToDevice));
    CheckOffsets<<<GridFor(n, 256, 1024), 256>>>(off, n, &d->max_deg, &d->bad);
    if (entries) CheckDests<<<GridFor(entries, 256, 2048), 256>>>(dst, entries, width, n, &d->bad);
    GG_HIP(hipGetLastError());
    GG_HIP(hipMemcpy(&h, d, sizeof(Flags), hipMemcpyDeviceToHost));
  } catch (...) {
    DeviceFree(d, sizeof(Flags));
    throw;
  }
  DeviceFree(d, sizeof(Flags));
  GG_REQUIRE(!h.bad, GG_EINVAL, "graph has decreasing offsets or a destination outside [0, n)");
  if (static_cast<int64_t>(h.max_deg) > *max_degree) *max_degree = static_cast<int64_t>(h.max_deg);
}

}  // namespace

bool GpuInitialised() { return g_initialised; }

void ValidateLists(const int64_t* off, const int32_t* dst, int64_t entries, int width, int64_t n,
                   int64_t* max_degree) {
  ValidateSide(off, dst, entries, width, n, max_degree);
}

gg_graph* NewDeviceGraph(int device, int64_t n, int64_t entries) {
  GG_REQUIRE(n >= 0 && n < (int64_t(1) << 31), GG_EINVAL, "num_nodes out of the int32 id range");
  gg_graph* g = new gg_graph;
  try {
    g->device = device;
    GG_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    GG_HIP(hipGetDeviceProperties(&prop, device));
    g->cu_count = prop.multiProcessorCount;
    GG_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    GG_HIP(hipEventCreate(&g->ev_start));
    GG_HIP(hipEventCreate(&g->ev_stop));
    g->n = n;
    g->out_entries = g->in_entries = entries;
    const size_t off_bytes = sizeof(int64_t) * (n + 1), dst_bytes = sizeof(int32_t) * entries;
    g->out_off = static_cast<int64_t*>(DeviceAlloc(off_bytes));
    g->out_dst = static_cast<int32_t*>(DeviceAlloc(dst_bytes));
    g->in_off = g->out_off;
    g->in_dst = g->out_dst;
    g->graph_bytes = off_bytes + dst_bytes;
    MailInit(g);
  } catch (...) {
    gg_graph_free(g);
    throw;
  }
  return g;
}

}  // namespace gg

using namespace gg;

extern "C" {

const char* gg_last_error(void) { return t_error.c_str(); }

int gg_init(int device) {
  return Guard([&] {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    GG_REQUIRE(e == hipSuccess && count > 0, GG_ENODEV, "no HIP device available");
    GG_REQUIRE(device >= 0 && device < count, GG_EINVAL, "device ordinal out of range");
    GG_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    GG_HIP(hipGetDeviceProperties(&prop, device));
    GG_REQUIRE(prop.warpSize == 64, GG_ENOTSUP, "kernels assume 64-wide wavefronts");
    NoAllocSetup(device);
    g_initialised = true;
  });
}

int gg_graph_upload(int64_t num_nodes, const int64_t* out_offsets, const int32_t* out_dests,
                    int weighted, int directed, const int64_t* in_offsets, const int32_t* in_dests,
                    gg_graph** out) {
  return Guard([&] {
    GG_REQUIRE(out != nullptr, GG_EINVAL, "null output handle");
    *out = nullptr;
    GG_REQUIRE(g_initialised, GG_ENODEV, "gg_init has not been called");
    GG_REQUIRE(num_nodes >= 0 && num_nodes < (int64_t(1) << 31), GG_EINVAL,
               "num_nodes out of the int32 id range");
    GG_REQUIRE(out_offsets != nullptr, GG_EINVAL, "null offsets");
    GG_REQUIRE(!directed || (in_offsets && (in_dests || in_offsets[num_nodes] == 0)), GG_EINVAL,
               "directed graphs need their inverse CSR");
    const int64_t e_out = out_offsets[num_nodes] - out_offsets[0];
    GG_REQUIRE(out_offsets[0] == 0 && e_out >= 0, GG_EINVAL, "offsets must start at 0");
    GG_REQUIRE(e_out == 0 || out_dests, GG_EINVAL, "null destinations");
    GG_REQUIRE(!directed || (in_offsets[0] == 0 && in_offsets[num_nodes] >= 0), GG_EINVAL,
               "inverse offsets must start at 0");
    const int width = weighted ? 2 : 1;

    gg_graph* g = new gg_graph;
    try {
      GG_HIP(hipGetDevice(&g->device));
      hipDeviceProp_t prop;
      GG_HIP(hipGetDeviceProperties(&prop, g->device));
      g->cu_count = prop.multiProcessorCount;
      GG_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
      GG_HIP(hipEventCreate(&g->ev_start));
      GG_HIP(hipEventCreate(&g->ev_stop));
      g->n = num_nodes;
      g->directed = directed != 0;
      g->weighted = weighted != 0;
      g->out_entries = e_out;
      const size_t off_bytes = sizeof(int64_t) * (num_nodes + 1);
      const size_t dst_bytes = sizeof(int32_t) * width * e_out;
      g->out_off = static_cast<int64_t*>(DeviceAlloc(off_bytes));
      g->out_dst = static_cast<int32_t*>(DeviceAlloc(dst_bytes));
      g->graph_bytes = off_bytes + dst_bytes;
      GG_HIP(hipMemcpy(g->out_off, out_offsets, off_bytes, hipMemcpyHostToDevice));
      if (dst_bytes) GG_HIP(hipMemcpy(g->out_dst, out_dests, dst_bytes, hipMemcpyHostToDevice));
      if (g->directed) {
        const int64_t e_in = in_offsets[num_nodes];
        const size_t idst_bytes = sizeof(int32_t) * width * e_in;
        g->in_entries = e_in;
        g->in_off = static_cast<int64_t*>(DeviceAlloc(off_bytes));
        g->in_dst = static_cast<int32_t*>(DeviceAlloc(idst_bytes));
        g->graph_bytes += off_bytes + idst_bytes;
        GG_HIP(hipMemcpy(g->in_off, in_offsets, off_bytes, hipMemcpyHostToDevice));
        if (idst_bytes) GG_HIP(hipMemcpy(g->in_dst, in_dests, idst_bytes, hipMemcpyHostToDevice));
      } else {
        g->in_entries = e_out;
        g->in_off = g->out_off;
        g->in_dst = g->out_dst;
      }
      GG_REQUIRE(out_offsets[num_nodes] == e_out, GG_EINVAL, "offsets[n] != entries");
      MailInit(g);
      ValidateSide(g->out_off, g->out_dst, g->out_entries, width, num_nodes, &g->max_degree);
      if (g->directed)
        ValidateSide(g->in_off, g->in_dst, g->in_entries, width, num_nodes, &g->max_degree);
    } catch (...) {
      gg_graph_free(g);
      throw;
    }
    *out = g;
  });
}

int gg_graph_copy_out(const gg_graph* g, int64_t* offsets, int32_t* dests) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_HIP(hipSetDevice(g->device));
    GG_HIP(hipStreamSynchronize(g->stream));
    const int width = g->weighted ? 2 : 1;
    if (offsets) GG_HIP(hipMemcpy(offsets, g->out_off, sizeof(int64_t) * (g->n + 1), hipMemcpyDeviceToHost));
    if (dests && g->out_entries)
      GG_HIP(hipMemcpy(dests, g->out_dst, sizeof(int32_t) * width * g->out_entries, hipMemcpyDeviceToHost));
  });
}

int gg_graph_free(gg_graph* g) {
  if (!g) return GG_OK;
  return Guard([&] {
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->ms_exec) (void)hipGraphExecDestroy(g->ms_exec);
    if (g->sb_exec) (void)hipGraphExecDestroy(g->sb_exec);
    g->ws.Release();
    const size_t off_bytes = sizeof(int64_t) * (g->n + 1);
    const int width = g->weighted ? 2 : 1;
    DeviceFree(g->out_off, off_bytes);
    DeviceFree(g->out_dst, sizeof(int32_t) * width * g->out_entries);
    if (g->directed) {
      DeviceFree(g->in_off, off_bytes);
      DeviceFree(g->in_dst, sizeof(int32_t) * width * g->in_entries);
    }
    DeviceFree(g->fold_dst, sizeof(int32_t) * g->out_entries);
    DeviceFree(g->new_id, sizeof(int32_t) * (g->n > 0 ? g->n : 1));
    DeviceFree(g->old_id, sizeof(int32_t) * (g->n > 0 ? g->n : 1));
    if (g->pinned) (void)hipHostFree(g->pinned);
    if (g->mail_host) (void)hipHostFree(const_cast<unsigned long long*>(g->mail_host));
    for (auto& p : g->pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : g->event_pool) (void)hipEventDestroy(e);
    if (g->side) (void)hipStreamSynchronize(g->side);
    if (g->side2) (void)hipStreamSynchronize(g->side2);
    if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
    if (g->ev_join) (void)hipEventDestroy(g->ev_join);
    if (g->ev_join2) (void)hipEventDestroy(g->ev_join2);
    if (g->side) (void)hipStreamDestroy(g->side);
    if (g->side2) (void)hipStreamDestroy(g->side2);
    if (g->ev_start) (void)hipEventDestroy(g->ev_start);
    if (g->ev_stop) (void)hipEventDestroy(g->ev_stop);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
  });
}

int64_t gg_graph_device_bytes(const gg_graph* g) {
  if (!g) return 0;
  return static_cast<int64_t>(g->graph_bytes + g->ws.Bytes());
}

int gg_profile_enable(gg_graph* g, int on) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    g->profile = on != 0;
  });
}

int gg_profile_reset(gg_graph* g) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_HIP(hipStreamSynchronize(g->stream));
    SettleProfile(g);
    g->kernel_totals.clear();
  });
}

int gg_profile_get(gg_graph* g, int idx, char* name, int name_len, int64_t* launches,
                   double* total_ms, double* alg_bytes) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(idx >= 0 && idx < static_cast<int>(g->kernel_totals.size()), GG_EINVAL,
               "profile index out of range");
    auto it = g->kernel_totals.begin();
    std::advance(it, idx);
    if (name && name_len > 0) std::snprintf(name, static_cast<size_t>(name_len), "%s", it->first.c_str());
    if (launches) *launches = it->second.launches;
    if (total_ms) *total_ms = it->second.ms;
    if (alg_bytes) *alg_bytes = it->second.bytes;
  });
}

}  // extern "C"
