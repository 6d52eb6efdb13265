// Afforest connected components on gfx950, and Shiloach-Vishkin
// (gg_cc_sv, src/cc_sv.cc:49-83) as a second variant.
//
// Reference: Afforest, Link, Compress, SampleFrequentElement
// (src/cc.cc:41-151).  Kept:
//   * Link hooks the higher root under the lower one with a compare-and-swap
//     and retries from the grandparents (cc.cc:41-55), so every component
//     ends labelled with its smallest vertex id -- the result is exactly the
//     reference's, not just the same partition;
//   * neighbor_rounds sampled rounds (default 2): each vertex links to its
//     r-th neighbour, then Compress (cc.cc:104-117);
//   * the most frequent label among 1024 samples drawn with a default-seeded
//     std::mt19937 and std::uniform_int_distribution over [0, N-1]
//     (cc.cc:69-92) names the component whose vertices skip the final link;
//   * the final link over neighbours from offset neighbor_rounds (plus all
//     in-neighbours for directed graphs) and a last Compress (cc.cc:120-150).
// GPU-specific: every read of comp[] inside Link and Compress is a relaxed
// agent-scope atomic load.  Plain loads may be served from a stale copy in
// another XCD's L2 during the kernel, and a retry loop that keeps reading a
// stale root would never finish; coherent loads keep the reference's
// progress guarantee.  The final link expands vertices over the grid with
// gg_expand.h.
#include <algorithm>
#include <mutex>
#include <random>

#include "gg_common.h"
#include "gg_expand.h"

// Fixed device scratch of gg_cc_noalloc, for callers that cannot allocate
// device memory.  Each GPU the library runs on holds its own copy, at its
// own address, from gg_init until the process ends.  In host code the name
// `scratch` is only a host-side stand-in: kernels get the device address
// that NoAllocSetup looked up (gg_noalloc_ptr[device]).
static constexpr size_t kNoAllocBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* gg_noalloc_ptr[kMaxDevices] = {};
}  // namespace

namespace gg {

void NoAllocSetup(int device) {
  GG_REQUIRE(device >= 0 && device < kMaxDevices, GG_ENOTSUP, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (gg_noalloc_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  GG_REQUIRE(e == hipSuccess && ptr != nullptr, GG_EHIP,
             std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  gg_noalloc_ptr[device] = ptr;
}

}  // namespace gg

namespace gg {
namespace {

constexpr int kBS = 256;
constexpr int kSamples = 1024;

// Labels are read and written at agent scope.  The MI355X's eight XCDs
// each have their own L2: a plain store can stay in the writer's L2 while
// another XCD reads the label from memory, and Shiloach-Vishkin takes an
// iteration without a hook as proof that every edge already joins equal
// labels.  With plain stores for its hooks and pointer jumps, 5 of 20 runs
// on a degree-ordered Kronecker 22 graph ended with a vertex still its own
// root beside 558 neighbours of label 0; with agent-scope stores, 0 of 20.
__device__ __forceinline__ int32_t Load(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void Store(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void Link(int32_t u, int32_t v, int32_t* comp) {
  int32_t p1 = Load(&comp[u]);
  int32_t p2 = Load(&comp[v]);
  while (p1 != p2) {
    const int32_t high = p1 > p2 ? p1 : p2;
    const int32_t low = p1 + (p2 - high);
    const int32_t p_high = Load(&comp[high]);
    if (p_high == low) break;
    if (p_high == high && atomicCAS(&comp[high], high, low) == high) break;
    p1 = Load(&comp[Load(&comp[high])]);
    p2 = Load(&comp[low]);
  }
}

__global__ __launch_bounds__(kBS) void CcInit(int32_t* __restrict__ comp, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
    comp[i] = static_cast<int32_t>(i);
}

// Sampled round r: link u with its r-th out-neighbour, if it has one.
__global__ __launch_bounds__(kBS) void CcLinkRound(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                   int32_t* comp, int64_t n, int r) {
  for (int64_t u = blockIdx.x * int64_t(kBS) + threadIdx.x; u < n; u += int64_t(gridDim.x) * kBS) {
    const int64_t b = off[u];
    if (off[u + 1] - b > r) Link(static_cast<int32_t>(u), dst[b + r], comp);
  }
}

// Pointer jumping until every vertex points at a root.
__global__ __launch_bounds__(kBS) void CcCompress(int32_t* comp, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS) {
    int32_t c = Load(&comp[i]);
    int32_t cc = Load(&comp[c]);
    while (c != cc) {
      c = cc;
      cc = Load(&comp[c]);
    }
    if (Load(&comp[i]) != c) Store(&comp[i], c);
  }
}

__global__ __launch_bounds__(kBS) void CcGather(const int32_t* __restrict__ comp, const int32_t* __restrict__ idx,
                                                int32_t* __restrict__ out, int k) {
  for (int i = blockIdx.x * kBS + threadIdx.x; i < k; i += gridDim.x * kBS) out[i] = comp[idx[i]];
}

struct LinkVisit {
  const int32_t* dst;
  int32_t* comp;
  __device__ __forceinline__ void operator()(int32_t u, int32_t, int64_t e, bool valid) {
    if (valid) Link(u, dst[e], comp);
  }
};

// Final link: every vertex outside the skipped component links with its
// neighbours from offset `start` on (start = 0 for the inverse CSR).
__global__ __launch_bounds__(kBS) void CcFinalLink(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                                   int32_t* comp, int64_t n, int32_t skip, int start,
                                                   EdgeChunk* chunks, int32_t* nchunks,
                                                   unsigned long long* examined) {
  __shared__ ExpandShared sh;
  __shared__ unsigned long long s_red[kBS / kWave];
  LinkVisit visit{dst, comp};
  unsigned long long ex = 0;
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < n; tile += int64_t(gridDim.x) * kBS) {
    const int64_t u = tile + threadIdx.x;
    int64_t beg = 0, end = 0;
    if (u < n && comp[u] != skip) {
      end = off[u + 1];
      beg = off[u] + start < end ? off[u] + start : end;
      ex += static_cast<unsigned long long>(end - beg);
    }
    ExpandTile(sh, static_cast<int32_t>(u), 0, beg, end, visit, chunks, nchunks, kChunkEdges);
  }
  ex = BlockSum<kBS>(ex, s_red);
  if (threadIdx.x == 0 && ex) atomicAdd(&examined[blockIdx.x & 63], ex);
}

__global__ __launch_bounds__(kBS) void CcFinalLinkChunks(const int32_t* __restrict__ dst, int32_t* comp,
                                                         const EdgeChunk* chunks, const int32_t* nchunks) {
  LinkVisit visit{dst, comp};
  ExpandChunks(chunks, *nchunks, visit);
}

// Shiloach-Vishkin's hook (src/cc_sv.cc:58-71): over edge (u, v), the higher
// of the two labels, while it is still a root, takes the lower one -- the
// min-max swap that makes lower labels win whatever the edge's direction.
// An agent-scope store where the reference has a plain one (see Load);
// `hooked` records that this thread changed a label.
struct HookVisit {
  const int32_t* dst;
  int32_t* comp;
  bool* hooked;
  __device__ __forceinline__ void operator()(int32_t u, int32_t, int64_t e, bool valid) {
    if (!valid) return;
    const int32_t cu = Load(&comp[u]), cv = Load(&comp[dst[e]]);
    if (cu == cv) return;
    const int32_t high = cu > cv ? cu : cv;
    const int32_t low = cu + (cv - high);
    if (Load(&comp[high]) == high) {
      Store(&comp[high], low);
      *hooked = true;
    }
  }
};

// One wavefront store per wavefront that hooked.
__device__ __forceinline__ void NoteChange(bool hooked, int* change) {
  if (__ballot(hooked) && LaneId() == 0) *change = 1;
}

__global__ __launch_bounds__(kBS) void SvHook(const int64_t* __restrict__ off, const int32_t* __restrict__ dst,
                                              int32_t* comp, int64_t n, EdgeChunk* chunks, int32_t* nchunks,
                                              int* change) {
  __shared__ ExpandShared sh;
  bool hooked = false;
  HookVisit visit{dst, comp, &hooked};
  for (int64_t tile = int64_t(blockIdx.x) * kBS; tile < n; tile += int64_t(gridDim.x) * kBS) {
    const int64_t u = tile + threadIdx.x;
    const int64_t beg = u < n ? off[u] : 0, end = u < n ? off[u + 1] : 0;
    ExpandTile(sh, static_cast<int32_t>(u), 0, beg, end, visit, chunks, nchunks, kChunkEdges);
  }
  NoteChange(hooked, change);
}

__global__ __launch_bounds__(kBS) void SvHookChunks(const int32_t* __restrict__ dst, int32_t* comp,
                                                    const EdgeChunk* chunks, const int32_t* nchunks, int* change) {
  bool hooked = false;
  HookVisit visit{dst, comp, &hooked};
  ExpandChunks(chunks, *nchunks, visit);
  NoteChange(hooked, change);
}

__global__ void SvClear(int* change, int32_t* nchunks) {
  if (threadIdx.x == 0) {
    *change = 0;
    *nchunks = 0;
  }
}

// The reference's sample positions: default-seeded mt19937 and
// uniform_int_distribution<int32_t>(0, n - 1), 1024 draws (cc.cc:75-80).
void SamplePositions(int64_t n, int32_t* idx) {
  std::mt19937 gen;
  std::uniform_int_distribution<int32_t> dist(0, static_cast<int32_t>(n - 1));
  for (int i = 0; i < kSamples; i++) idx[i] = dist(gen);
}

// Clears the examined counters and the chunk count (a kernel, not a memset,
// so the sequence holds no memset node if a caller captures it).
__global__ void CcClear(unsigned long long* examined, int32_t* nchunks) {
  if (threadIdx.x < 64) examined[threadIdx.x] = 0;
  if (threadIdx.x == 0) nchunks[0] = 0;
}

__global__ void CcClearChunks(int32_t* nchunks) {
  if (threadIdx.x == 0) nchunks[0] = 0;
}

// The most frequent label among the samples; ties go to the smallest label
// (the choice only decides which component skips the final link, never the
// labels).  Sorts the samples in place.
int32_t MostFrequent(int32_t* s, int k) {
  std::sort(s, s + k);
  int32_t best_label = s[0];
  int best = 0;
  for (int i = 0; i < k;) {
    int j = i;
    while (j < k && s[j] == s[i]) j++;
    if (j - i > best) {
      best = j - i;
      best_label = s[i];
    }
    i = j;
  }
  return best_label;
}

// Device buffers of one Afforest run.
struct CcBuffers {
  int32_t* comp;                 // n
  int32_t* sample_idx;           // kSamples, filled by the caller
  int32_t* sample_val;           // kSamples
  unsigned long long* examined;  // 64 striped counters
  int32_t* nchunks;              // 1
  EdgeChunk* chunks;             // MaxChunks(n, entries), or null when no vertex reaches kChunkEdges
};

// Afforest (cc.cc:95-151) on g->stream into b.comp, inside the caller's
// timed region.  h_samples (host, kSamples) receives the sampled labels.
// Returns the kernel launches issued.
int64_t RunAfforest(gg_graph* g, int neighbor_rounds, const CcBuffers& b, int32_t* h_samples) {
  const int64_t n = g->n;
  hipStream_t st = g->stream;
  const dim3 blk(kBS);
  const dim3 grid_n(GridFor(n, kBS, g->cu_count * 32));
  const double N = static_cast<double>(n);
  int64_t launches = 0;
  hipLaunchKernelGGL(CcClear, dim3(1), dim3(64), 0, st, b.examined, b.nchunks);
  Run(g, "cc_init", 4 * N, CcInit, grid_n, blk, b.comp, n);
  launches += 2;
  for (int r = 0; r < neighbor_rounds; r++) {
    // offsets 16 B (both ends), the r-th neighbour, comp read/write.
    Run(g, "cc_link_round", 16 * N + 4 * N + 8 * N, CcLinkRound, grid_n, blk,
        static_cast<const int64_t*>(g->out_off), static_cast<const int32_t*>(g->out_dst), b.comp, n, r);
    Run(g, "cc_compress", 8 * N, CcCompress, grid_n, blk, b.comp, n);
    launches += 2;
  }
  CcGather<<<dim3(kSamples / kBS), blk, 0, st>>>(b.comp, b.sample_idx, b.sample_val, kSamples);
  GG_HIP(hipMemcpyAsync(h_samples, b.sample_val, sizeof(int32_t) * kSamples, hipMemcpyDeviceToHost, st));
  GG_HIP(hipStreamSynchronize(st));
  launches++;
  const int32_t skip = MostFrequent(h_samples, kSamples);
  auto final_link = [&](const int64_t* off, const int32_t* dst, int start) {
    hipLaunchKernelGGL(CcClearChunks, dim3(1), dim3(64), 0, st, b.nchunks);
    Run(g, "cc_final_link", 8 * N + 4.0 * (off == g->out_off ? g->out_entries : g->in_entries), CcFinalLink,
        dim3(GridFor(n, kBS, g->cu_count * 8)), blk, off, dst, b.comp, n, skip, start, b.chunks, b.nchunks,
        b.examined);
    launches += 2;
    if (b.chunks) {
      Run(g, "cc_final_link_chunks", 0, CcFinalLinkChunks, dim3(g->cu_count * 4), blk, dst, b.comp,
          static_cast<const EdgeChunk*>(b.chunks), static_cast<const int32_t*>(b.nchunks));
      launches++;
    }
  };
  final_link(g->out_off, g->out_dst, neighbor_rounds);
  if (g->directed) final_link(g->in_off, g->in_dst, 0);
  Run(g, "cc_compress", 8 * N, CcCompress, grid_n, blk, b.comp, n);
  launches++;
  GG_HIP(hipGetLastError());
  return launches;
}

int64_t CcEntries(const gg_graph* g) { return g->out_entries > g->in_entries ? g->out_entries : g->in_entries; }

void CcCheck(const gg_graph* g, int neighbor_rounds) {
  GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
  GG_REQUIRE(!g->weighted, GG_ENOTSUP, "CC needs an unweighted graph upload");
  GG_REQUIRE(neighbor_rounds >= 0, GG_EINVAL, "neighbor_rounds must be >= 0");
  GG_REQUIRE(g->n > 0, GG_EINVAL, "empty graph");
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_cc(gg_graph* g, int neighbor_rounds, int32_t* comp_out, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    CcCheck(g, neighbor_rounds);
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n;
    CcBuffers b;
    b.comp = g->Scratch<int32_t>("cc.comp", n);
    b.sample_idx = g->Scratch<int32_t>("cc.sample_idx", kSamples);
    b.sample_val = g->Scratch<int32_t>("cc.sample_val", kSamples);
    b.examined = g->Scratch<unsigned long long>("cc.examined", 64);
    b.nchunks = g->Scratch<int32_t>("cc.nchunks", 2);
    b.chunks = g->max_degree >= kChunkEdges ? g->Scratch<EdgeChunk>("cc.chunks", MaxChunks(n, CcEntries(g)))
                                            : nullptr;
    int32_t* h_samples = g->Pinned<int32_t>(kSamples);
    if (g->cc_samples_n != n) {  // sample positions depend only on n
      int32_t idx[kSamples];
      SamplePositions(n, idx);
      GG_HIP(hipMemcpy(b.sample_idx, idx, sizeof(int32_t) * kSamples, hipMemcpyHostToDevice));
      g->cc_samples_n = n;
    }
    BeginTiming(g);
    const int64_t launches = RunAfforest(g, neighbor_rounds, b, h_samples);
    FinishTiming(g, stats, wt);
    if (comp_out) CopyOutComponents(g, b.comp, comp_out);
    if (stats) {
      unsigned long long h[64];
      GG_HIP(hipMemcpy(h, b.examined, sizeof(h), hipMemcpyDeviceToHost));
      unsigned long long ex = 0;
      for (unsigned long long x : h) ex += x;
      stats->total_ms = wt.ms();
      stats->iterations = neighbor_rounds + 1;
      stats->edges_examined = static_cast<int64_t>(ex);
      stats->launches = launches;
    }
  });
}

// Allocation-free Afforest: gg_cc's kernels with every device buffer carved
// out of the fixed `scratch` array of the current GPU (no hipMalloc, no
// hipFree, no pinned or library allocation).  Layout, each part 256-byte
// aligned: comp[n], the sample positions and labels, the striped counters,
// the chunk count, the hub chunk list, and for a degree-ordered handle the
// two per-vertex arrays of the label mapping.
extern "C" int gg_cc_noalloc(gg_graph* g, int neighbor_rounds, int32_t* comp_out, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    CcCheck(g, neighbor_rounds);
    GG_HIP(hipSetDevice(g->device));
    void* base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_noalloc_mu);
      if (g->device >= 0 && g->device < kMaxDevices) base = gg_noalloc_ptr[g->device];
    }
    GG_REQUIRE(base != nullptr, GG_ENODEV, "gg_init has not been called for this graph's device");
    const int64_t n = g->n;
    const bool use_chunks = g->max_degree >= kChunkEdges;
    size_t used = 0;
    auto carve = [&](size_t bytes) {
      const size_t at = used;
      used += (bytes + 255) & ~size_t(255);
      return at;
    };
    const size_t at_comp = carve(sizeof(int32_t) * n);
    const size_t at_idx = carve(sizeof(int32_t) * kSamples);
    const size_t at_val = carve(sizeof(int32_t) * kSamples);
    const size_t at_ex = carve(sizeof(unsigned long long) * 64);
    const size_t at_nc = carve(sizeof(int32_t) * 2);
    const size_t at_chunks = use_chunks ? carve(sizeof(EdgeChunk) * MaxChunks(n, CcEntries(g))) : 0;
    const size_t at_least = g->new_id ? carve(sizeof(int32_t) * n) : 0;
    const size_t at_out = g->new_id ? carve(sizeof(int32_t) * n) : 0;
    GG_REQUIRE(used <= kNoAllocBytes, GG_EINVAL,
               "graph needs " + std::to_string(used) + " bytes of scratch; gg_cc_noalloc holds " +
                   std::to_string(kNoAllocBytes) + " (use gg_cc)");
    unsigned char* p = static_cast<unsigned char*>(base);
    CcBuffers b;
    b.comp = reinterpret_cast<int32_t*>(p + at_comp);
    b.sample_idx = reinterpret_cast<int32_t*>(p + at_idx);
    b.sample_val = reinterpret_cast<int32_t*>(p + at_val);
    b.examined = reinterpret_cast<unsigned long long*>(p + at_ex);
    b.nchunks = reinterpret_cast<int32_t*>(p + at_nc);
    b.chunks = use_chunks ? reinterpret_cast<EdgeChunk*>(p + at_chunks) : nullptr;
    // The scratch keeps whatever the last call on this GPU left, so the
    // sample positions are written every call (RunAfforest clears the rest).
    int32_t idx[kSamples], h_samples[kSamples];
    SamplePositions(n, idx);
    GG_HIP(hipMemcpyAsync(b.sample_idx, idx, sizeof(idx), hipMemcpyHostToDevice, g->stream));
    BeginTiming(g);
    const int64_t launches = RunAfforest(g, neighbor_rounds, b, h_samples);
    FinishTiming(g, stats, wt);
    if (comp_out) {
      if (g->new_id)
        CopyOutComponentsWith(g, b.comp, comp_out, reinterpret_cast<int32_t*>(p + at_least),
                              reinterpret_cast<int32_t*>(p + at_out));
      else
        CopyOutComponentsWith(g, b.comp, comp_out, nullptr, nullptr);
    }
    if (stats) {
      unsigned long long h[64];
      GG_HIP(hipMemcpy(h, b.examined, sizeof(h), hipMemcpyDeviceToHost));
      unsigned long long ex = 0;
      for (unsigned long long x : h) ex += x;
      stats->total_ms = wt.ms();
      stats->iterations = neighbor_rounds + 1;
      stats->edges_examined = static_cast<int64_t>(ex);
      stats->launches = launches;
    }
  });
}

extern "C" int gg_cc_sv(gg_graph* g, int32_t* comp_out, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE(!g->weighted, GG_ENOTSUP, "CC needs an unweighted graph upload");
    GG_REQUIRE(g->n > 0, GG_EINVAL, "empty graph");
    GG_HIP(hipSetDevice(g->device));
    const int64_t n = g->n;
    hipStream_t st = g->stream;
    const dim3 blk(kBS), grid_n(GridFor(n, kBS, g->cu_count * 32));
    const double N = static_cast<double>(n);
    int32_t* comp = g->Scratch<int32_t>("cc.comp", n);
    int32_t* nchunks = g->Scratch<int32_t>("cc.nchunks", 2);
    int* change = g->Scratch<int>("sv.change", 1);
    const bool use_chunks = g->max_degree >= kChunkEdges;
    EdgeChunk* chunks = use_chunks ? g->Scratch<EdgeChunk>("cc.chunks", MaxChunks(n, g->out_entries)) : nullptr;
    int* h_change = g->Pinned<int>(1);
    int64_t iterations = 0, launches = 0;
    BeginTiming(g);
    Run(g, "cc_init", 4 * N, CcInit, grid_n, blk, comp, n);
    launches++;
    // cc_sv.cc:55-80: hook over every out-edge, compress, until no hook.
    for (;;) {
      iterations++;
      Run(g, "sv_clear", 0, SvClear, dim3(1), dim3(64), change, nchunks);
      Run(g, "sv_hook", 16 * N + 4.0 * g->out_entries, SvHook, dim3(GridFor(n, kBS, g->cu_count * 8)), blk,
          static_cast<const int64_t*>(g->out_off), static_cast<const int32_t*>(g->out_dst), comp, n,
          use_chunks ? chunks : static_cast<EdgeChunk*>(nullptr), nchunks, change);
      if (use_chunks)
        Run(g, "sv_hook_chunks", 0, SvHookChunks, dim3(g->cu_count * 4), blk,
            static_cast<const int32_t*>(g->out_dst), comp, static_cast<const EdgeChunk*>(chunks),
            static_cast<const int32_t*>(nchunks), change);
      Run(g, "cc_compress", 8 * N, CcCompress, grid_n, blk, comp, n);
      launches += use_chunks ? 4 : 3;
      GG_HIP(hipMemcpyAsync(h_change, change, sizeof(int), hipMemcpyDeviceToHost, st));
      GG_HIP(hipStreamSynchronize(st));
      if (*h_change == 0) break;
      GG_REQUIRE(iterations < n + 2, GG_EHIP, "Shiloach-Vishkin did not converge");
    }
    GG_HIP(hipGetLastError());
    FinishTiming(g, stats, wt);
    if (comp_out) CopyOutComponents(g, comp, comp_out);
    if (stats) {
      stats->total_ms = wt.ms();
      stats->iterations = iterations;
      stats->edges_examined = iterations * g->out_entries;
      stats->launches = launches;
    }
  });
}
