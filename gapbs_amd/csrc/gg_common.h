// Shared pieces of the gfx950 kernel library: error handling, the device
// graph handle with its cached workspaces, and wave64 primitives.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "gapgpu.h"

namespace gg {

// ---------------------------------------------------------------------------
// Errors: every C-ABI entry point runs its body inside Guard(), which turns
// these exceptions into a negative status and a thread-local message.
struct Failure : std::runtime_error {
  int status;
  Failure(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void SetError(const std::string& msg);

#define GG_HIP(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::gg::Failure(e_ == hipErrorOutOfMemory ? GG_ENOMEM : GG_EHIP,                 \
                          std::string(#expr) + ": " + hipGetErrorString(e_) + " at " +     \
                              __FILE__ + ":" + std::to_string(__LINE__));                  \
  } while (0)

#define GG_REQUIRE(cond, status, msg)                  \
  do {                                                 \
    if (!(cond)) throw ::gg::Failure((status), (msg)); \
  } while (0)

template <class F>
int Guard(F&& f) {
  try {
    f();
    return GG_OK;
  } catch (const Failure& e) {
    SetError(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    SetError("host allocation failed");
    return GG_ENOMEM;
  } catch (const std::exception& e) {
    SetError(e.what());
    return GG_EHIP;
  }
}

// Device allocation with the optional GAPGPU_MEM_CAP_BYTES cap (lets tests
// drive the out-of-memory path from the host).
void* DeviceAlloc(size_t bytes);
void DeviceFree(void* p, size_t bytes);

// ---------------------------------------------------------------------------
// Device graph handle.
struct Workspace {
  struct Slot {
    void* p = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Slot> slots;
  // Returns a buffer of at least `bytes`, reallocating if the cached one is
  // smaller.  Contents are undefined.
  void* Get(const std::string& name, size_t bytes);
  void Release();
  // Frees the buffers whose names start with `prefix`.
  void Drop(const std::string& prefix);
  size_t Bytes() const;
};

}  // namespace gg

struct gg_graph {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  hipStream_t side = nullptr, side2 = nullptr;  // EnsureSideStream
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
  int64_t n = 0;
  int64_t out_entries = 0, in_entries = 0;
  bool directed = false, weighted = false;
  int64_t* out_off = nullptr;
  int32_t* out_dst = nullptr;  // int32 ids, or {v, w} pairs when weighted
  int64_t* in_off = nullptr;   // == out_off for undirected graphs
  int32_t* in_dst = nullptr;
  size_t graph_bytes = 0;
  // Degree-ordered layout (gg_graph_order_by_degree): the lists above use
  // device ids, new_id[original id] is a vertex's device id and
  // old_id[device id] its original id.  Both null: ids as uploaded.
  int32_t* new_id = nullptr;
  int32_t* old_id = nullptr;
  // Degree-ordered unweighted handle: the lists in uploaded order (each row
  // sorted by uploaded id, entries in device ids), built on first use by
  // UploadOrderLists.  Null otherwise.
  int32_t* fold_dst = nullptr;
  int64_t max_degree = 0;  // max out-degree (in-degree for the inverse), set at upload
  int64_t no_in_count = -1;  // vertices without in-neighbours (computed on first BFS)
  gg_tuning tune = {15, 18, 15, 32, 32, 1024, 16, 0, {0, 0, 0, 0}, {0, 0}, {0, 0}, 0, 0, 0};  // gg_autotune results
  int64_t pr_bins[3] = {-1, -1, -1};     // PageRank in-degree bins (vertex counts)
  int64_t pr_bin_edges[3] = {0, 0, 0};   // and their in-edge counts
  // Jacobi PageRank source-segment layout, one per precision (gg_pr.hip):
  // the in-entries whose source lies below `hot`, regrouped by source
  // segment of 2^sb vertices into (destination, sub-list) pairs; the device
  // arrays live in the workspace ("pr.seg32.*" / "pr.seg64.*").
  struct PrSegments {
    bool built = false;
    int sb = 0;                              // log2 width of the L2 segments
    int sbl = 0;                             // log2 width of the LDS segments (-1: none)
    int n_lds = 0;                           // LDS segments (the first n_lds of ns)
    int32_t small_max = 0, medium_max = 0;   // pair bin cut-offs (sub-list lengths)
    int32_t min_density = 0;                 // and the density rule
    int ns = 0;                              // segments in use (0: plain pull)
    int64_t hot = 0;                         // sources below this go by segment
    int64_t pairs = 0, hot_entries = 0;
    std::vector<int64_t> first_pair;         // ns + 1: segment s owns pairs [first_pair[s], first_pair[s + 1])
    std::vector<int64_t> large, medium;      // its pairs in the block and wavefront bins (listed first)
    std::vector<int64_t> entries;            // its in-entries
    double build_ms = 0;
  } pr_seg[2];
  // Jacobi PageRank propagation-block layout (gg_pr.hip), shared by both
  // precisions: in-entries grouped by (source tile, destination bin) of 2^lg
  // vertices each; device arrays in the workspace ("pr.pb.*").
  struct PrBlocks {
    bool built = false;
    int lg = 0;                       // log2 of the tile / bin width; 0: no blocks
    int64_t k = 0;                    // tiles (= bins)
    int64_t octets = 0;               // padded entries / 8
    int64_t n_write = 0, n_sum = 0;   // chunks of the write and sum phases
    int64_t n_multi = 0, slots = 0;   // bins summed in several chunks, their partial-sum slots
    double build_ms = 0;
  } pr_blocks;
  int64_t cc_samples_n = -1;             // n the uploaded Afforest sample positions are for
  int64_t ms_heavy = -1;                 // batched BFS: heavy vertices (listed once per graph)
  int64_t ms_heavy_chunks = 0;           // and their in-edge chunks
  int64_t mb_hubs = -1;                  // batched BC: vertices whose dependency folds are split
  int64_t mb_big_hubs = -1;              // (those below and above the wide-block cut-off)
  hipGraphExec_t ms_exec = nullptr;      // batched traversal: two captured levels, replayed
  std::vector<uintptr_t> ms_exec_key;    // the buffers and arguments captured in it
  int ms_exec_kernels = 0;               // kernels per replay
  std::vector<int64_t> pk_parts;         // packed traversal: the part list the lists were last checked against
  hipGraphExec_t sb_exec = nullptr;      // batched SSSP: captured rounds, replayed
  std::vector<uintptr_t> sb_exec_key;    // the buffers and arguments captured in it
  gg::Workspace ws;
  void* pinned = nullptr;  // small pinned host block for counter read-back
  size_t pinned_bytes = 0;
  // Mailbox: pinned, device-mapped words the GPU posts small per-step
  // results into; word 0 is a sequence number the host polls (gg::MailWait).
  volatile unsigned long long* mail_host = nullptr;
  unsigned long long* mail_dev = nullptr;
  unsigned long long mail_seq = 0;
  int cu_count = 256;

  // Per-kernel profiling (gg_profile_*): launches bracketed by events.
  struct KernelTotals {
    int64_t launches = 0;
    double ms = 0;
    double bytes = 0;
  };
  struct PendingLaunch {
    std::string name;
    hipEvent_t a, b;
    double bytes;
  };
  bool profile = false;
  std::map<std::string, KernelTotals> kernel_totals;
  std::vector<PendingLaunch> pending;
  std::vector<hipEvent_t> event_pool;

  template <class T>
  T* Scratch(const char* name, size_t count) {
    return static_cast<T*>(ws.Get(name, count * sizeof(T)));
  }
  template <class T>
  T* Pinned(size_t count);
};

namespace gg {

// True once gg_init has run in this process.
bool GpuInitialised();
// Checks device lists (offsets non-decreasing, destinations in [0, n); width
// 2 = {v, w} pairs) and raises *max_degree to their largest degree; throws
// GG_EINVAL otherwise.
void ValidateLists(const int64_t* off, const int32_t* dst, int64_t entries, int width, int64_t n,
                   int64_t* max_degree);
// A new undirected, unweighted handle on `device` with uninitialised device
// arrays for n vertices and `entries` destinations (filled by the caller).
gg_graph* NewDeviceGraph(int device, int64_t n, int64_t entries);

// Vertex ids across the C-ABI are the uploaded ones; a degree-ordered graph
// maps them on the way in and its per-vertex results on the way out.
// DeviceVertex: the device id of original vertex v (synchronous read).
int32_t DeviceVertex(gg_graph* g, int32_t v);
// In place on g->stream: k original ids at d (device memory) -> device ids.
void DeviceVertices(gg_graph* g, int32_t* d, int64_t k);
// Copies a per-vertex array in device order (src, device memory) to host in
// original order; with ids, non-negative values are vertex ids and are
// mapped too.  Synchronous.
void CopyOutVertexI32(gg_graph* g, const int32_t* src, int32_t* host, bool ids);
void CopyOutVertexF32(gg_graph* g, const float* src, float* host);
// Component labels in device order, each the smallest device id of its
// component -> host, each the smallest original id of its component.
void CopyOutComponents(gg_graph* g, const int32_t* comp, int32_t* host);
// The same with the caller's two device work arrays of n int32 (unused, and
// may be null, when the handle is not degree-ordered).
void CopyOutComponentsWith(gg_graph* g, const int32_t* comp, int32_t* host, int32_t* least, int32_t* tmp);
// The out-lists in uploaded order: g->out_dst unless the handle is
// degree-ordered and unweighted (its rows are sorted by device id), then
// g->fold_dst, built here on first use (device ids, each row sorted by
// uploaded id).  Kernels whose result depends on the order in which a row
// is summed read these.  Synchronous on first use; call outside timing.
const int32_t* UploadOrderLists(gg_graph* g);
// Looks up, once per device, the device address of gg_cc_noalloc's fixed
// scratch array on `device` (the current device); called by gg_init.
void NoAllocSetup(int device);
// Propagation blocks (gg_pr.hip): the in-entries grouped by (source tile,
// destination bin) of 2^lg vertices, stored tile-major in octets of 8
// entries (16-bit offsets within the tile and the bin; padding entries
// have destination kPbPad), with a bin-major octet index and chunk tables.
constexpr uint32_t kPbPad = 0xFFFFu;
struct PbChunk {
  uint32_t part;    // the tile (write phase) or bin (sum phase)
  int32_t slot;     // sum phase: partial-result slot, -1 when the chunk is the whole bin
  uint32_t q0, q1;  // write: octets [q0, q1); sum: octet-index positions [q0, q1)
};
struct PbMulti {
  uint32_t bin, slot0, nslots, pad;  // a bin summed in nslots chunks, slots slot0...
};
struct PbArrays {
  const uint4* src16;        // per octet: 8 source offsets within the tile
  const uint4* dst16;        // per octet: 8 destination offsets within the bin
  const uint32_t* oct;       // each bin's octets, bin-major
  const PbChunk* wchunks;    // write-phase chunks (PrBlocks::n_write)
  const PbChunk* schunks;    // sum-phase chunks (PrBlocks::n_sum)
  const PbMulti* multi;      // bins of several sum chunks (PrBlocks::n_multi)
};
// The layout FP32 Jacobi PageRank runs through on g, built on first use
// (null when it runs without blocks), and its device arrays.
const gg_graph::PrBlocks* PropagationBlocks(gg_graph* g);
PbArrays PropagationBlockArrays(gg_graph* g);

// Direction-optimizing BFS (gg_bfs.hip) from device id `source` on
// g->stream, inside the caller's timed region: parents in the workspace
// array "bfs.parent", depth (device, may be null) the level of every vertex
// (-1 unreached).  Also what BC's discovery uses: depths are unique, so any
// BFS gives the reference's levels.
struct BfsRun {
  int64_t td = 0, bu = 0;  // levels run in each direction
  int32_t levels = 0;      // deepest level + 1 ... as counted by the level loop
  int64_t reached = 0, reach = 0, examined = 0, vexamined = 0, launches = 0;
};
BfsRun RunDobfs(gg_graph* g, int32_t source, int alpha, int beta, int32_t* depth);

// Host wall clock for gg_stats::total_ms.
struct WallTimer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

inline void StatsReset(gg_stats* s) {
  if (s) std::memset(s, 0, sizeof(*s));
}

// Records the start event; FinishTiming records stop, waits for the stream,
// fills device_ms and settles pending profiled launches.
void BeginTiming(gg_graph* g);
void FinishTiming(gg_graph* g, gg_stats* s, const WallTimer& wt);

// Profiling: CreditBytes adds algorithmic bytes known only after a kernel
// ran to its record.
void CreditBytes(gg_graph* g, int idx, double bytes);
// Folds finished launches into the per-kernel totals (stream must be idle).
void SettleProfile(gg_graph* g);

// Takes two events from the pool (creating them on demand) and records a
// pending profiled launch; returns its index.
int ProfileSlot(gg_graph* g, const char* name, double bytes, hipEvent_t* a, hipEvent_t* b);

// Launches kernel k on g->stream.  When profiling is on, the dispatch packet
// itself records start/stop events (hipExtLaunchKernelGGL: no extra queue
// packets), and the launch is credited to `name` with `bytes` algorithmic
// bytes; returns the record index for CreditBytes, else -1.
template <class... KArgs, class... Args>
int RunOn(gg_graph* g, hipStream_t stream, const char* name, double bytes, void (*k)(KArgs...), dim3 grid,
          dim3 block, Args... args) {
  if (!g->profile) {
    hipLaunchKernelGGL(k, grid, block, 0, stream, args...);
    return -1;
  }
  hipEvent_t a, b;
  const int idx = ProfileSlot(g, name, bytes, &a, &b);
  hipExtLaunchKernelGGL(k, grid, block, 0, stream, a, b, 0, args...);
  return idx;
}

template <class... KArgs, class... Args>
int Run(gg_graph* g, const char* name, double bytes, void (*k)(KArgs...), dim3 grid, dim3 block,
        Args... args) {
  return RunOn(g, g->stream, name, bytes, k, grid, block, args...);
}

// Two more streams of the handle (default priority, created on first use)
// for kernels that may overlap the main stream's, with fork/join events.
void EnsureSideStream(gg_graph* g);

// Mailbox protocol.  A one-block kernel calls MailPost with the next
// sequence number after writing up to kMailWords values; the host waits for
// it with MailWait, which spins briefly and otherwise falls back to a
// stream synchronisation (so a lost post can never hang the host).
constexpr int kMailWords = 32;
void MailInit(gg_graph* g);
// Returns the sequence number the next MailPost must carry.
inline unsigned long long MailNext(gg_graph* g) { return ++g->mail_seq; }
// Waits until the mailbox holds `seq`; copies values[0 .. k) out.
void MailWait(gg_graph* g, unsigned long long seq, unsigned long long* values, int k);

// Exact work of a kernel kind, counted for the byte credits of a profiled
// run: kWorkStripes striped counters of list entries read, then as many of
// matched entries (the entries whose neighbour the kernel went on to read).
constexpr int kWorkStripes = 64;
constexpr int kWorkCounters = 2 * kWorkStripes;

namespace {
__global__ __attribute__((unused)) void WorkClear(unsigned long long* work, int count) {
  for (int i = threadIdx.x; i < count; i += blockDim.x) work[i] = 0;
}
}  // namespace
// Sums the counters of `kinds` consecutive kinds: {entries, matched} each.
std::vector<unsigned long long> ReadWork(gg_graph* g, const unsigned long long* work, int kinds);

inline unsigned GridFor(int64_t items, int block, unsigned cap = 1u << 20) {
  int64_t b = (items + block - 1) / block;
  if (b < 1) b = 1;
  if (b > static_cast<int64_t>(cap)) b = cap;
  return static_cast<unsigned>(b);
}

// ---------------------------------------------------------------------------
// wave64 device primitives.
constexpr int kWave = 64;

__device__ __forceinline__ int LaneId() { return __lane_id(); }

__device__ __forceinline__ uint64_t LanesBelow(uint64_t mask) {
  return mask & ((uint64_t(1) << LaneId()) - 1);
}

// Wave-aggregated append: every active lane calls it together; lanes with
// `take` get consecutive slots in `out` after one atomic per wave.
template <class T>
__device__ __forceinline__ void WaveAppend(bool take, T value, T* out, int32_t* tail) {
  const uint64_t m = __ballot(take);
  if (m == 0) return;
  const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  int base = 0;
  if (LaneId() == leader) base = atomicAdd(tail, __popcll(m));
  base = __shfl(base, leader);
  if (take) out[base + __popcll(LanesBelow(m))] = value;
}

template <class T>
__device__ __forceinline__ T WaveSum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Inclusive scan within a wave.
__device__ __forceinline__ int WaveInclusiveScan(int v) {
  const int lane = LaneId();
  for (int o = 1; o < kWave; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one int per thread; returns the prefix and
// writes the block total.  `sh` needs BS/64 ints of shared memory.
template <int BS>
__device__ __forceinline__ int BlockExclusiveScan(int v, int* total, int* sh) {
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  const int inc = WaveInclusiveScan(v);
  if (lane == kWave - 1) sh[wid] = inc;
  __syncthreads();
  int wave_base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BS / kWave; w++) {
    const int s = sh[w];
    if (w < wid) wave_base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_base + inc - v;
}

// Adds a block's totals to its stripe; call from every thread with values
// valid in thread 0 (e.g. after BlockSum).
__device__ __forceinline__ void WorkAdd(unsigned long long* work, unsigned long long entries,
                                        unsigned long long matched) {
  if (threadIdx.x != 0) return;
  const int s = blockIdx.x & (kWorkStripes - 1);
  if (entries) atomicAdd(&work[s], entries);
  if (matched) atomicAdd(&work[kWorkStripes + s], matched);
}

// Writes values to the host mailbox, then the sequence number, with system
// scope so the host sees the values before the sequence number.  Call from
// one thread.
__device__ __forceinline__ void MailPost(unsigned long long* mb, const unsigned long long* values,
                                         int k, unsigned long long seq) {
  for (int i = 0; i < k; i++)
    __hip_atomic_store(&mb[1 + i], values[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __threadfence_system();
  __hip_atomic_store(&mb[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Block-wide sum, result valid in every thread.  `sh` needs BS/64 slots.
template <int BS, class T>
__device__ __forceinline__ T BlockSum(T v, T* sh) {
  v = WaveSum(v);
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  T t = 0;
#pragma unroll
  for (int w = 0; w < BS / kWave; w++) t += sh[w];
  __syncthreads();
  return t;
}

}  // namespace gg

template <class T>
T* gg_graph::Pinned(size_t count) {
  const size_t bytes = count * sizeof(T);
  if (bytes > pinned_bytes) {
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_bytes = 0;
    GG_HIP(hipHostMalloc(&pinned, bytes < 4096 ? 4096 : bytes, hipHostMallocDefault));
    pinned_bytes = bytes < 4096 ? 4096 : bytes;
  }
  return static_cast<T*>(pinned);
}
