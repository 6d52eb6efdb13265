// Shared plumbing of the command-line drivers: each driver is the
// reference's main() (src/<kernel>.cc) with the kernel call replaced by a
// call through the gapgpu C-ABI.  The graph is uploaded to HBM lazily, on
// the first trial, so `-n0` runs (the reference's generate/load smoke tests)
// never touch the GPU.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <memory>

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
  DeviceSession(int device, const GraphT& g) : device_(device), g_(g) {}
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
    }
    return dg_;
  }

 private:
  int device_;
  const GraphT& g_;
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

inline void LogStats(bool on, const gg_stats& st) {
  if (!on) return;
  PrintStep("device ms", st.device_ms);
  PrintStep("steps", st.iterations);
}

}  // namespace gapdrv
