// converter: the reference's converter (src/converter.cc:14-31) -- builds
// or loads a graph and writes it as .sg/.wsg (-b) or as a text edge list
// (-e).  Host only; no GPU.
#include <fstream>

#include "driver_common.h"

using namespace gapdrv;

template <class GraphT>
void WriteEdgeList(const GraphT& g, const std::string& path) {
  if (path.empty()) throw Error(-8, "No output filename given (Use -h for help)");
  std::ofstream out(path);
  if (!out) throw Error(-5, "Couldn't write to file " + path);
  for (NodeID u = 0; u < g.num_nodes(); u++)
    for (const auto& d : g.out_neigh(u)) {
      out << u << " " << DestId(d);
      if (std::is_same<typename GraphT::Dest, WNode>::value)
        out << " " << reinterpret_cast<const WNode&>(d).w;
      out << "\n";
    }
}

template <class DestT>
void Convert(const CLI& cli) {
  CSRGraphT<DestT> g = MakeGraph<DestT>(cli);
  g.PrintStats();
  if (cli.out_sg())
    WriteSerialized(g, cli.out_filename());
  else
    WriteEdgeList(g, cli.out_filename());
}

int main(int argc, char* argv[]) {
  return RunDriver([&] {
    CLI cli(argc, argv, "converter", CLI::kConvert);
    cli.ParseArgs();
    if (cli.out_weighted())
      Convert<WNode>(cli);
    else
      Convert<NodeID>(cli);
    return 0;
  });
}
