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
// (hipMemcpyAsync on the handle's stream), kReaders host threads read the
// next chunk of the file into the other with pread, a quarter each; an
// event per buffer says when it may be refilled.  The host never holds more
// than the two buffers.  The lists are then
// checked on the device as an upload's are (offsets from 0 to entries and
// non-decreasing, destinations in range).
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "gg_common.h"

namespace gg {
namespace {

constexpr size_t kStage = size_t(256) << 20;
constexpr int kReaders = 4;

struct Staging {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int fd = -1;
  int64_t pos = 0;  // file offset of the next byte to stream
  ~Staging() {
    for (int i = 0; i < 2; i++) {
      if (ev[i]) {
        (void)hipEventSynchronize(ev[i]);
        (void)hipEventDestroy(ev[i]);
      }
      if (buf[i]) (void)hipHostFree(buf[i]);
    }
    if (fd >= 0) close(fd);
  }
};

// Reads exactly `bytes` at file offset `at` into p (false on a short read).
bool ReadAt(int fd, char* p, size_t bytes, int64_t at) {
  while (bytes) {
    const ssize_t r = pread(fd, p, bytes, static_cast<off_t>(at));
    if (r <= 0) return false;
    p += r;
    bytes -= static_cast<size_t>(r);
    at += r;
  }
  return true;
}

// The next `bytes` of the file into device memory dev, chunk by chunk.
void StreamInto(Staging& s, hipStream_t st, void* dev, size_t bytes, int* next, const std::string& path) {
  size_t done = 0;
  while (done < bytes) {
    const int b = *next;
    const size_t chunk = bytes - done < kStage ? bytes - done : kStage;
    GG_HIP(hipEventSynchronize(s.ev[b]));  // its previous copy has finished
    const size_t part = (chunk + kReaders - 1) / kReaders;
    bool ok[kReaders];
    std::vector<std::thread> readers;
    for (int t = 0; t < kReaders; t++) {
      const size_t lo = t * part, len = lo < chunk ? (chunk - lo < part ? chunk - lo : part) : 0;
      ok[t] = true;
      if (len)
        readers.emplace_back([&, t, lo, len] {
          ok[t] = ReadAt(s.fd, static_cast<char*>(s.buf[b]) + lo, len, s.pos + static_cast<int64_t>(lo));
        });
    }
    for (auto& r : readers) r.join();
    for (int t = 0; t < kReaders; t++) GG_REQUIRE(ok[t], GG_EINVAL, "Short read from " + path);
    GG_HIP(hipMemcpyAsync(static_cast<char*>(dev) + done, s.buf[b], chunk, hipMemcpyHostToDevice, st));
    GG_HIP(hipEventRecord(s.ev[b], st));
    s.pos += static_cast<int64_t>(chunk);
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
    s.fd = open(path, O_RDONLY);
    GG_REQUIRE(s.fd >= 0, GG_EINVAL, "Couldn't open file " + p);
    uint8_t directed = 0;
    int64_t entries = 0, n = 0;
    GG_REQUIRE(ReadAt(s.fd, reinterpret_cast<char*>(&directed), 1, 0) &&
                   ReadAt(s.fd, reinterpret_cast<char*>(&entries), 8, 1) &&
                   ReadAt(s.fd, reinterpret_cast<char*>(&n), 8, 9),
               GG_EINVAL, "Short read from " + p);
    s.pos = 17;
    GG_REQUIRE(n >= 0 && n < (int64_t(1) << 31) && entries >= 0, GG_EINVAL, "Corrupt header in " + p);
    // Entry counts are bounded (2^40 entries is 8 TB of lists) before any
    // size is computed, so the byte counts below cannot wrap; the file must
    // then hold exactly the header and the lists it announces.
    GG_REQUIRE(entries < (int64_t(1) << 40), GG_EINVAL, "Corrupt header in " + p);
    const int width = weighted ? 2 : 1;
    const size_t off_bytes = sizeof(int64_t) * (n + 1), dst_bytes = sizeof(int32_t) * width * entries;
    struct stat fs;
    GG_REQUIRE(fstat(s.fd, &fs) == 0, GG_EINVAL, "Couldn't stat file " + p);
    const uint64_t want = 17 + (directed ? 2 : 1) * (uint64_t(off_bytes) + dst_bytes);
    GG_REQUIRE(static_cast<uint64_t>(fs.st_size) >= want, GG_EINVAL, "Short read from " + p);
    GG_REQUIRE(static_cast<uint64_t>(fs.st_size) == want, GG_EINVAL, "Corrupt header in " + p);
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
