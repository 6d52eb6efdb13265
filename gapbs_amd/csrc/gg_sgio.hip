// gg_graph_read: a .sg / .wsg file straight into HBM.
//
// Reference: Reader::ReadSerializedGraph (src/reader.h:254-307), in the
// layout the host reader has (gap_sgio.h) and the reference writes
// (src/writer.h:39-69):
//   bool directed (1 byte) | int64 entries | int64 nodes |
//   int64 offsets[nodes + 1] | dest[entries]   (int32, or {int32 v, int32 w})
//   and, for directed graphs, the inverse offsets and dests in the same form.
// The file streams through two pinned staging buffers (hipHostMalloc):
// while the copy engine moves one buffer to its place in HBM
// (hipMemcpyAsync on the handle's stream), the host reads the next chunk of
// the file into the other; an event per buffer says when it may be refilled.
// The host never holds more than the two buffers.  The lists are then
// checked on the device as an upload's are (offsets from 0 to entries and
// non-decreasing, destinations in range).
#include <cstdio>
#include <string>

#include "gg_common.h"

namespace gg {
namespace {

constexpr size_t kStage = size_t(64) << 20;

struct Staging {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  FILE* f = nullptr;
  ~Staging() {
    for (int i = 0; i < 2; i++) {
      if (ev[i]) {
        (void)hipEventSynchronize(ev[i]);
        (void)hipEventDestroy(ev[i]);
      }
      if (buf[i]) (void)hipHostFree(buf[i]);
    }
    if (f) std::fclose(f);
  }
};

// bytes from the file into device memory dev, chunk by chunk.
void StreamInto(Staging& s, hipStream_t st, void* dev, size_t bytes, int* next, const std::string& path) {
  size_t done = 0;
  while (done < bytes) {
    const int b = *next;
    const size_t chunk = bytes - done < kStage ? bytes - done : kStage;
    GG_HIP(hipEventSynchronize(s.ev[b]));  // its previous copy has finished
    GG_REQUIRE(std::fread(s.buf[b], 1, chunk, s.f) == chunk, GG_EINVAL, "Short read from " + path);
    GG_HIP(hipMemcpyAsync(static_cast<char*>(dev) + done, s.buf[b], chunk, hipMemcpyHostToDevice, st));
    GG_HIP(hipEventRecord(s.ev[b], st));
    done += chunk;
    *next = b ^ 1;
  }
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_graph_read(const char* path, int weighted, gg_graph** out, int64_t* num_nodes, gg_stats* stats) {
  return Guard([&] {
    WallTimer wt;
    StatsReset(stats);
    GG_REQUIRE(out != nullptr, GG_EINVAL, "null output handle");
    *out = nullptr;
    GG_REQUIRE(GpuInitialised(), GG_ENODEV, "gg_init has not been called");
    GG_REQUIRE(path != nullptr, GG_EINVAL, "null path");
    const std::string p(path);
    const size_t dot = p.rfind('.');
    GG_REQUIRE(dot != std::string::npos, GG_EINVAL, "Couldn't find suffix of " + p);
    const std::string suffix = p.substr(dot);
    GG_REQUIRE(suffix == ".sg" || suffix == ".wsg", GG_EINVAL, "Unrecognized suffix: " + suffix);
    GG_REQUIRE(weighted || suffix != ".wsg", GG_EINVAL, ".wsg only allowed for weighted graphs");
    GG_REQUIRE(!weighted || suffix != ".sg", GG_EINVAL, ".sg not allowed for weighted graphs");
    Staging s;
    s.f = std::fopen(path, "rb");
    GG_REQUIRE(s.f != nullptr, GG_EINVAL, "Couldn't open file " + p);
    uint8_t directed = 0;
    int64_t entries = 0, n = 0;
    GG_REQUIRE(std::fread(&directed, 1, 1, s.f) == 1 && std::fread(&entries, 8, 1, s.f) == 1 &&
                   std::fread(&n, 8, 1, s.f) == 1,
               GG_EINVAL, "Short read from " + p);
    GG_REQUIRE(n >= 0 && n < (int64_t(1) << 31) && entries >= 0, GG_EINVAL, "Corrupt header in " + p);
    const int width = weighted ? 2 : 1;
    const size_t off_bytes = sizeof(int64_t) * (n + 1), dst_bytes = sizeof(int32_t) * width * entries;
    int device = 0;
    GG_HIP(hipGetDevice(&device));
    for (int i = 0; i < 2; i++) {
      GG_HIP(hipHostMalloc(&s.buf[i], kStage, hipHostMallocDefault));
      GG_HIP(hipEventCreateWithFlags(&s.ev[i], hipEventDisableTiming));
    }
    gg_graph* g = NewDeviceGraph(device, n, 0);
    try {
      g->weighted = weighted != 0;
      DeviceFree(g->out_dst, 0);
      g->out_dst = g->in_dst = nullptr;
      g->out_dst = static_cast<int32_t*>(DeviceAlloc(dst_bytes));
      g->out_entries = g->in_entries = entries;
      g->in_dst = g->out_dst;
      g->graph_bytes = off_bytes + dst_bytes;
      int next = 0;
      StreamInto(s, g->stream, g->out_off, off_bytes, &next, p);
      StreamInto(s, g->stream, g->out_dst, dst_bytes, &next, p);
      if (directed) {
        g->in_off = static_cast<int64_t*>(DeviceAlloc(off_bytes));
        g->in_dst = nullptr;
        g->directed = true;
        g->in_dst = static_cast<int32_t*>(DeviceAlloc(dst_bytes));
        g->graph_bytes += off_bytes + dst_bytes;
        StreamInto(s, g->stream, g->in_off, off_bytes, &next, p);
        StreamInto(s, g->stream, g->in_dst, dst_bytes, &next, p);
      }
      GG_HIP(hipStreamSynchronize(g->stream));
      const int sides = directed ? 2 : 1;
      for (int side = 0; side < sides; side++) {
        const int64_t* off = side ? g->in_off : g->out_off;
        int64_t ends[2] = {0, 0};
        GG_HIP(hipMemcpy(&ends[0], off, sizeof(int64_t), hipMemcpyDeviceToHost));
        GG_HIP(hipMemcpy(&ends[1], off + n, sizeof(int64_t), hipMemcpyDeviceToHost));
        GG_REQUIRE(ends[0] == 0 && ends[1] == entries, GG_EINVAL, "Corrupt offsets in " + p);
        ValidateLists(off, side ? g->in_dst : g->out_dst, entries, width, n, &g->max_degree);
      }
    } catch (...) {
      gg_graph_free(g);
      throw;
    }
    *out = g;
    if (num_nodes) *num_nodes = n;
    if (stats) {
      stats->total_ms = wt.ms();
      stats->reached_edges = entries;
    }
  });
}
