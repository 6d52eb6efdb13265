// Serialized graph files (.sg unweighted, .wsg weighted).
//
// Reference: Reader::ReadSerializedGraph (src/reader.h:254-307) and
// WriterBase::WriteSerializedGraph (src/writer.h:39-69).  The byte layout is
// kept identical so that graphs move between these drivers and the
// reference binaries (`-f file.sg`):
//   bool directed (1 byte) | int64 entries | int64 nodes |
//   int64 offsets[nodes+1] | dest[entries]            (dest: int32 or {int32,int32})
//   and, for directed graphs, the inverse offsets and dests in the same form.
#pragma once

#include <cstdio>
#include <string>

#include "gap_graph.h"

namespace gap {

inline std::string FileSuffix(const std::string& path) {
  const size_t p = path.rfind('.');
  if (p == std::string::npos) throw Error(-1, "Couldn't find suffix of " + path);
  return path.substr(p);
}

namespace detail {
struct File {
  FILE* f = nullptr;
  File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};
inline void ReadExact(FILE* f, void* p, size_t bytes, const std::string& path) {
  if (bytes && std::fread(p, 1, bytes, f) != bytes) throw Error(-6, "Short read from " + path);
}
inline void WriteExact(FILE* f, const void* p, size_t bytes, const std::string& path) {
  if (bytes && std::fwrite(p, 1, bytes, f) != bytes) throw Error(-5, "Couldn't write to file " + path);
}
}  // namespace detail

template <class DestT>
CSRGraphT<DestT> ReadSerialized(const std::string& path) {
  constexpr bool kWeighted = std::is_same<DestT, WNode>::value;
  const std::string suffix = FileSuffix(path);
  if (suffix != ".sg" && suffix != ".wsg") throw Error(-3, "Unrecognized suffix: " + suffix);
  if (!kWeighted && suffix == ".wsg") throw Error(-5, ".wsg only allowed for weighted graphs");
  if (kWeighted && suffix == ".sg") throw Error(-5, ".sg not allowed for weighted graphs");
  detail::File file(path, "rb");
  if (!file.f) throw Error(-6, "Couldn't open file " + path);
  uint8_t directed = 0;
  int64_t entries = 0, n = 0;
  detail::ReadExact(file.f, &directed, 1, path);
  detail::ReadExact(file.f, &entries, 8, path);
  detail::ReadExact(file.f, &n, 8, path);
  if (n < 0 || entries < 0) throw Error(-6, "Corrupt header in " + path);
  auto read_side = [&](Buf<Offset>* off, Buf<DestT>* dst) {
    off->reset(static_cast<size_t>(n) + 1);
    dst->reset(static_cast<size_t>(entries));
    detail::ReadExact(file.f, off->data(), sizeof(Offset) * (n + 1), path);
    detail::ReadExact(file.f, dst->data(), sizeof(DestT) * entries, path);
  };
  Buf<Offset> off, ioff;
  Buf<DestT> dst, idst;
  read_side(&off, &dst);
  if (!directed) return CSRGraphT<DestT>(n, std::move(off), std::move(dst));
  read_side(&ioff, &idst);
  return CSRGraphT<DestT>(n, std::move(off), std::move(dst), std::move(ioff), std::move(idst));
}

template <class DestT>
void WriteSerialized(const CSRGraphT<DestT>& g, const std::string& path) {
  if (path.empty()) throw Error(-8, "No output filename given (Use -h for help)");
  detail::File file(path, "wb");
  if (!file.f) throw Error(-5, "Couldn't write to file " + path);
  const uint8_t directed = g.directed() ? 1 : 0;
  const int64_t entries = g.num_edges_directed();
  const int64_t n = g.num_nodes();
  detail::WriteExact(file.f, &directed, 1, path);
  detail::WriteExact(file.f, &entries, 8, path);
  detail::WriteExact(file.f, &n, 8, path);
  detail::WriteExact(file.f, g.out_offsets(), sizeof(Offset) * (n + 1), path);
  detail::WriteExact(file.f, g.out_dests(), sizeof(DestT) * entries, path);
  if (directed) {
    detail::WriteExact(file.f, g.in_offsets(), sizeof(Offset) * (n + 1), path);
    detail::WriteExact(file.f, g.in_dests(), sizeof(DestT) * entries, path);
  }
}

}  // namespace gap
