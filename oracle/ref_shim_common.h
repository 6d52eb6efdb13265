// Test infrastructure only: shared part of the shims that expose the
// reference's own kernels (compiled from the reference sources) to the
// parity tests and to the CPU baseline.  Included by ref_shim_<kernel>.cc
// after the reference translation unit, so the reference's CSRGraph,
// pvector and SourcePicker types are in scope.
//
// A graph crosses the shim as plain int64 offsets and int32 (or {int32 v,
// int32 w}) neighbour arrays -- the .sg/.wsg layout (src/writer.h:39-69).
#pragma once

#include <chrono>
#include <cstring>

namespace refshim {

template <class DestT>
DestT* CopyDests(const void* src, int64_t entries) {
  DestT* d = new DestT[entries > 0 ? entries : 1];
  if (entries > 0) std::memcpy(d, src, sizeof(DestT) * entries);
  return d;
}

template <class DestT>
CSRGraph<NodeID, DestT>* MakeGraph(int64_t n, const int64_t* off, const void* dst, int directed,
                                   const int64_t* ioff, const void* idst) {
  pvector<SGOffset> offsets(n + 1);
  for (int64_t i = 0; i <= n; i++) offsets[i] = off[i];
  DestT* neighs = CopyDests<DestT>(dst, off[n]);
  DestT** index = CSRGraph<NodeID, DestT>::GenIndex(offsets, neighs);
  if (!directed) return new CSRGraph<NodeID, DestT>(n, index, neighs);
  pvector<SGOffset> in_offsets(n + 1);
  for (int64_t i = 0; i <= n; i++) in_offsets[i] = ioff[i];
  DestT* in_neighs = CopyDests<DestT>(idst, ioff[n]);
  DestT** in_index = CSRGraph<NodeID, DestT>::GenIndex(in_offsets, in_neighs);
  return new CSRGraph<NodeID, DestT>(n, index, neighs, in_index, in_neighs);
}

inline double Seconds(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

}  // namespace refshim
