// Shared plumbing of the command-line drivers: each driver is the
// reference's main() (src/<kernel>.cc) with the kernel call replaced by a
// call through the gapgpu C-ABI.  The graph is uploaded to HBM lazily, on
// the first trial, so `-n0` runs (the reference's generate/load smoke tests)
// never touch the GPU.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <memory>
#include <vector>

#include "gap_build.h"
#include "gap_core.h"
#include "gap_graph.h"
#include "gap_harness.h"
#include "gap_verify.h"
#include "gapgpu.h"

namespace gapdrv {

using namespace gap;

// A GPU failure ends the driver with the library's status, like the
// reference's fatal errors end with their exit code.
[[noreturn]] inline void GpuFail(int rc, const char* what) {
  std::cout << what << " failed (" << rc << "): " << gg_last_error() << std::endl;
  std::exit(rc);
}

inline void Check(int rc, const char* what) {
  if (rc < 0) GpuFail(rc, what);
}

template <class GraphT>
class DeviceSession {
 public:
  // ordered: renumber the device copy by degree after the upload (-o);
  // every call still takes and returns the graph's own ids.
  DeviceSession(int device, const GraphT& g, bool ordered = false) : device_(device), g_(g), ordered_(ordered) {}
  ~DeviceSession() {
    if (dg_) gg_graph_free(dg_);
  }
  gg_graph* get() {
    if (!dg_) {
      Check(gg_init(device_), "gg_init");
      constexpr bool kWeighted = std::is_same<typename GraphT::Dest, WNode>::value;
      Check(gg_graph_upload(g_.num_nodes(), g_.out_offsets(),
                            reinterpret_cast<const int32_t*>(g_.out_dests()), kWeighted,
                            g_.directed(), g_.directed() ? g_.in_offsets() : nullptr,
                            g_.directed() ? reinterpret_cast<const int32_t*>(g_.in_dests()) : nullptr,
                            &dg_),
            "gg_graph_upload");
      if (ordered_) {
        if (g_.directed()) {
          std::cout << "Degree ordering skipped: directed graph" << std::endl;
        } else {
          gg_stats st;
          Check(gg_graph_order_by_degree(dg_, &st), "gg_graph_order_by_degree");
          gap::PrintTime("Degree Ordering", st.device_ms / 1e3);
        }
      }
    }
    return dg_;
  }

 private:
  int device_;
  const GraphT& g_;
  bool ordered_;
  gg_graph* dg_ = nullptr;
};

// Runs a driver body; host errors exit with the reference's status codes.
inline int RunDriver(const std::function<int()>& body) {
  try {
    return body();
  } catch (const gap::Error& e) {
    std::cout << e.what() << std::endl;
    std::exit(e.code);
  }
}

// -q: tunes the selected kernels (GG_TUNE_* bits) on this graph before
// any trial, from the first k roots the trials will use (a SourcePicker of
// its own with the same seed), and prints the choice.
template <class GraphT>
gg_tuning Autotune(const CLI& cli, DeviceSession<GraphT>& dev, const GraphT& g, int what, int k) {
  std::vector<int32_t> roots;
  if (what != GG_TUNE_PR) {
    SourcePicker<GraphT> sp(g, static_cast<NodeID>(cli.start_vertex()));
    for (int i = 0; i < k; i++) roots.push_back(sp.PickNext());
  }
  gg_tuning t;
  Check(gg_autotune(dev.get(), what, roots.empty() ? nullptr : roots.data(), static_cast<int>(roots.size()), &t),
        "gg_autotune");
  if (what & GG_TUNE_BATCH)
    std::cout << "Autotune: batched alpha " << t.batch_alpha << " (" << t.best_ms[1] << " ms)" << std::endl;
  if (what & GG_TUNE_BFS)
    std::cout << "Autotune: alpha " << t.bfs_alpha << " beta " << t.bfs_beta << " (" << t.best_ms[0] << " ms)"
              << std::endl;
  if (what & GG_TUNE_SSSP)
    std::cout << "Autotune: delta " << t.sssp_delta << " (" << t.best_ms[2] << " ms)" << std::endl;
  if (what & GG_TUNE_PR)
    std::cout << "Autotune: bins <" << t.pr_small_max << " <" << t.pr_medium_max << ", " << t.pr_blocks_per_cu
              << " blocks/CU (" << t.best_ms[3] << " ms/iter)" << std::endl;
  return t;
}

inline void LogStats(bool on, const gg_stats& st) {
  if (!on) return;
  PrintStep("device ms", st.device_ms);
  PrintStep("steps", st.iterations);
}

}  // namespace gapdrv
