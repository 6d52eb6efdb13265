// Device-wide primitives for the graph rebuilds: an exclusive prefix sum and
// a stable LSD radix sort of 64-bit keys.
//
// Radix sort: 8-bit digits over the bit range the caller names.  A pass is
// three launches:
//   RsHistogram  per-tile digit counts, stored digit-major (digit x tiles);
//   ScanU32      exclusive scan of those counts: the output position of
//                each (digit, tile) run;
//   RsScatter    each tile sorts its keys by the digit in LDS with eight
//                stable 1-bit splits (block prefix sums), then writes every
//                key to its run's position + its rank inside the tile's run.
// Tiles are kRsThreads x kRsItems = 4096 keys.
#pragma once

#include "gg_common.h"

namespace gg {
namespace {

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one value per thread (kRsThreads threads);
// *total gets the sum.  sh needs kRsThreads / 64 slots.
template <class T>
__device__ __forceinline__ T BlockScanExcl(T v, T* total, T* sh) {
  const int lane = LaneId(), wid = threadIdx.x / kWave;
  T inc = v;
  for (int o = 1; o < kWave; o <<= 1) {
    const T t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == kWave - 1) sh[wid] = inc;
  __syncthreads();
  T base = 0, tot = 0;
  for (int w = 0; w < kRsThreads / kWave; w++) {
    const T s = sh[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// ---- exclusive prefix sum (in place), T = uint32_t or int64_t ---------------

// Scans each tile in place and stores the tile's total.
template <class T>
__global__ __launch_bounds__(kScanThreads) void ScanTiles(T* __restrict__ data, int64_t n, T* __restrict__ sums) {
  __shared__ T sh[kScanThreads / kWave];
  const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
  T v[kScanItems];
  T local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = base + i < n ? data[base + i] : T(0);
    local += v[i];
  }
  T total;
  T run = BlockScanExcl<T>(local, &total, sh);
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < n) data[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the tile totals by one block (tiles <= a few 10^5).
template <class T>
__global__ __launch_bounds__(kScanThreads) void ScanSums(T* __restrict__ sums, int64_t tiles, T* grand) {
  __shared__ T sh[kScanThreads / kWave];
  const int64_t per = (tiles + kScanThreads - 1) / kScanThreads;
  const int64_t lo = per * threadIdx.x;
  const int64_t hi = lo + per < tiles ? lo + per : tiles;
  T local = 0;
  for (int64_t i = lo; i < hi; i++) local += sums[i];
  T total;
  T run = BlockScanExcl<T>(local, &total, sh);
  for (int64_t i = lo; i < hi; i++) {
    const T s = sums[i];
    sums[i] = run;
    run += s;
  }
  if (threadIdx.x == 0 && grand) *grand = total;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void ScanAdd(T* __restrict__ data, int64_t n, const T* __restrict__ sums) {
  const int64_t base = int64_t(blockIdx.x) * kScanTile;
  const T add = sums[blockIdx.x];
  for (int i = threadIdx.x; i < kScanTile; i += kScanThreads)
    if (base + i < n) data[base + i] += add;
}

// In-place exclusive scan of data[0, n); *grand (device, nullable) gets the
// total.  Workspace names are prefixed with `tag`.
template <class T>
void DeviceScan(gg_graph* g, const char* tag, T* data, int64_t n, T* grand) {
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  T* sums = g->Scratch<T>((std::string(tag) + ".sums").c_str(), tiles + 1);
  const int64_t t = tiles > 0 ? tiles : 1;
  Run(g, "scan", 0, ScanTiles<T>, dim3(static_cast<unsigned>(t)), dim3(kScanThreads), data, n, sums);
  Run(g, "scan", 0, ScanSums<T>, dim3(1), dim3(kScanThreads), sums, tiles, grand);
  Run(g, "scan", 0, ScanAdd<T>, dim3(static_cast<unsigned>(t)), dim3(kScanThreads), data, n,
      static_cast<const T*>(sums));
}

// ---- radix sort ---------------------------------------------------------------

__global__ __launch_bounds__(kRsThreads) void RsHistogram(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                          int64_t tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = tile * kRsTile;
    for (int i = threadIdx.x; i < kRsTile; i += kRsThreads)
      if (base + i < n) atomicAdd(&h[(keys[base + i] >> shift) & 255], 1u);
    __syncthreads();
    hist[int64_t(threadIdx.x) * tiles + tile] = h[threadIdx.x];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kRsThreads) void RsScatter(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                        int64_t n, int shift, int64_t tiles,
                                                        const uint32_t* __restrict__ offsets) {
  __shared__ uint64_t a[kRsTile];
  __shared__ uint32_t first[256];
  __shared__ uint32_t sh[kRsThreads / kWave];
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t base = tile * kRsTile;
    const int count = static_cast<int>(n - base < kRsTile ? n - base : kRsTile);
    // Missing keys of the last tile sort last (digit 256 in effect).
    for (int i = threadIdx.x; i < kRsTile; i += kRsThreads) a[i] = i < count ? in[base + i] : ~0ull;
    __syncthreads();
    // Eight stable 1-bit splits on bits shift .. shift+7; thread t owns the
    // blocked range [t * kRsItems, (t + 1) * kRsItems).
    for (int b = 0; b < 8; b++) {
      uint64_t v[kRsItems];
      uint32_t zeros = 0;
#pragma unroll
      for (int i = 0; i < kRsItems; i++) {
        const int l = threadIdx.x * kRsItems + i;
        v[i] = a[l];
        const bool pad = l >= count;
        zeros += !pad && !((v[i] >> (shift + b)) & 1);
      }
      uint32_t total_zeros;
      uint32_t zrank = BlockScanExcl<uint32_t>(zeros, &total_zeros, sh);
      // Ones (and padding) go after all zeros, in order.
      uint32_t orank = total_zeros + (threadIdx.x * kRsItems - zrank);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kRsItems; i++) {
        const int l = threadIdx.x * kRsItems + i;
        const bool zero = l < count && !((v[i] >> (shift + b)) & 1);
        if (zero)
          a[zrank++] = v[i];
        else
          a[orank++] = v[i];
      }
      __syncthreads();
    }
    // First local index of each digit (the tile is now sorted by digit).
    first[threadIdx.x] = 0xFFFFFFFFu;
    __syncthreads();
    for (int l = threadIdx.x; l < count; l += kRsThreads) {
      const uint32_t d = (a[l] >> shift) & 255;
      if (l == 0 || ((a[l - 1] >> shift) & 255) != d) first[d] = l;
    }
    __syncthreads();
    for (int l = threadIdx.x; l < count; l += kRsThreads) {
      const uint32_t d = (a[l] >> shift) & 255;
      out[offsets[int64_t(d) * tiles + tile] + (l - first[d])] = a[l];
    }
    __syncthreads();
  }
}

// Stable sort of keys[0, n) by bits [lo_bit, hi_bit) in 8-bit digits from
// lo_bit up (a digit reaching past hi_bit must be constant there); calls on
// increasing bit ranges compose into one LSD sort.  tmp holds n keys.
// Returns the buffer with the result (keys or tmp).  n < 2^32.
inline uint64_t* RadixSort(gg_graph* g, const char* tag, uint64_t* keys, uint64_t* tmp, int64_t n, int lo_bit,
                           int hi_bit) {
  GG_REQUIRE(n < (int64_t(1) << 32), GG_ENOTSUP, "radix sort of 2^32 keys or more");
  if (n <= 1) return keys;
  const int64_t tiles = (n + kRsTile - 1) / kRsTile;
  uint32_t* hist = g->Scratch<uint32_t>((std::string(tag) + ".hist").c_str(), tiles * 256);
  const dim3 grid(GridFor(tiles, 1, g->cu_count * 8)), blk(kRsThreads);
  for (int shift = lo_bit; shift < hi_bit; shift += 8) {
    Run(g, "radix_sort", 0, RsHistogram, grid, blk, static_cast<const uint64_t*>(keys), n, shift, tiles, hist);
    DeviceScan<uint32_t>(g, (std::string(tag) + ".scan").c_str(), hist, tiles * 256, nullptr);
    Run(g, "radix_sort", 16.0 * n, RsScatter, grid, blk, static_cast<const uint64_t*>(keys), tmp, n, shift, tiles,
        static_cast<const uint32_t*>(hist));
    std::swap(keys, tmp);
  }
  return keys;
}

}  // namespace
}  // namespace gg
