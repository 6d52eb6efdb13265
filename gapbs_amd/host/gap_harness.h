// Command line, graph acquisition, root picking and the trial harness.
//
// Reference counterparts:
//   CLBase / CLApp / CLIterApp / CLPageRank / CLDelta ... src/command_line.h:28-226
//   BuilderBase::MakeGraph ............................. src/builder.h:343-364
//   SourcePicker ....................................... src/benchmark.h:46-68
//   TopK ............................................... src/benchmark.h:72-88
//   BenchmarkKernel .................................... src/benchmark.h:98-123
//
// The reference's flags keep their letters and defaults.  Flags added for
// the GPU path use letters the reference leaves free:
//   -z <seed>   generator seed (default 27491095, the reference's only seed)
//   -c <side>   generate the side x side weighted 2D grid graph
//   -x <dev>    HIP device ordinal
//   -j <k>      roots per batched launch (bfs, sssp, bc; default 1 = per root)
//   -p <mode>   score precision: pr fp32 (compensated, default) | gs | fp64;
//               bc fp32 (the reference's float arithmetic, default) | kahan | fp64
//   -q          autotune direction/bucket/bin parameters for this graph
#pragma once

#include <getopt.h>

#include <algorithm>
#include <functional>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "gap_build.h"
#include "gap_core.h"
#include "gap_gen.h"
#include "gap_graph.h"
#include "gap_sgio.h"

namespace gap {

class CLI {
 public:
  enum Kind { kBase = 0, kApp = 1, kIter = 2, kPageRank = 3, kDelta = 4, kConvert = 5 };

  CLI(int argc, char** argv, std::string name, Kind kind, int default_iters = 1,
      double default_tol = 1e-4)
      : argc_(argc), argv_(argv), name_(std::move(name)), kind_(kind) {
    opts_ = "f:g:hk:su:mz:c:x:";
    Help('h', "", "print this help message");
    Help('f', "file", "load graph from file");
    Help('s', "", "symmetrize input edge list", "false");
    Help('g', "scale", "generate 2^scale kronecker graph");
    Help('u', "scale", "generate 2^scale uniform-random graph");
    Help('k', "degree", "average degree for synthetic graph", std::to_string(degree_));
    Help('m', "", "reduces memory usage during graph building", "false");
    Help('z', "seed", "seed of the synthetic graph generator", std::to_string(seed_));
    Help('c', "side", "generate side x side weighted 2D grid graph");
    Help('x', "dev", "HIP device ordinal", "0");
    if (kind == kConvert) {
      opts_ += "e:b:w";
      Help('b', "file", "output serialized graph to file");
      Help('e', "file", "output edge list to file");
      Help('w', "", "make output weighted");
      return;
    }
    if (kind == kBase) return;
    opts_ += "an:r:vlj:p:qo";
    Help('a', "", "output analysis of last run", "false");
    Help('n', "n", "perform n trials", std::to_string(num_trials_));
    Help('r', "node", "start from node r", "rand");
    Help('v', "", "verify the output of each run", "false");
    Help('l', "", "log performance within each trial", "false");
    Help('j', "k", "roots per batched GPU launch", "1");
    Help('p', "mode", "score precision: fp32, kahan (bc), gs (pr) or fp64", "fp32");
    Help('q', "", "autotune GPU parameters for this graph", "false");
    Help('o', "", "renumber the GPU copy by degree (results keep ids)", "false");
    num_iters_ = default_iters;
    max_iters_ = default_iters;
    tolerance_ = default_tol;
    if (kind == kIter) {
      opts_ += "i:";
      Help('i', "i", "perform i iterations", std::to_string(num_iters_));
    } else if (kind == kPageRank) {
      opts_ += "i:t:";
      Help('i', "i", "perform at most i iterations", std::to_string(max_iters_));
      Help('t', "t", "use tolerance t", std::to_string(tolerance_));
    } else if (kind == kDelta) {
      opts_ += "d:";
      Help('d', "d", "delta parameter", std::to_string(delta_));
    }
  }

  bool ParseArgs() {
    optind = 1;
    int c;
    while ((c = getopt(argc_, argv_, opts_.c_str())) != -1) Handle(c, optarg);
    if (filename_.empty() && scale_ == -1 && grid_side_ == -1) {
      std::cout << "No graph input specified. (Use -h for help)" << std::endl;
      return false;
    }
    if (scale_ != -1 || grid_side_ != -1) symmetrize_ = true;
    return true;
  }

  void PrintUsage() const {
    std::cout << name_ << std::endl;
    for (const auto& h : help_) std::cout << h << std::endl;
    std::exit(0);
  }

  int scale() const { return scale_; }
  int degree() const { return degree_; }
  const std::string& filename() const { return filename_; }
  bool symmetrize() const { return symmetrize_; }
  bool uniform() const { return uniform_; }
  bool in_place() const { return in_place_; }
  int64_t seed() const { return seed_; }
  int64_t grid_side() const { return grid_side_; }
  int device() const { return device_; }
  bool do_analysis() const { return do_analysis_; }
  int num_trials() const { return num_trials_; }
  int64_t start_vertex() const { return start_vertex_; }
  bool do_verify() const { return do_verify_; }
  bool logging_en() const { return logging_; }
  int batch() const { return batch_; }
  const std::string& precision() const { return precision_; }
  bool autotune() const { return autotune_; }
  bool order_by_degree() const { return order_by_degree_; }
  int num_iters() const { return num_iters_; }
  int max_iters() const { return max_iters_; }
  double tolerance() const { return tolerance_; }
  int32_t delta() const { return delta_; }
  const std::string& out_filename() const { return out_filename_; }
  bool out_weighted() const { return out_weighted_; }
  bool out_el() const { return out_el_; }
  bool out_sg() const { return out_sg_; }

 private:
  void Help(char opt, std::string arg, const std::string& text, std::string def = "") {
    char buf[128];
    if (!arg.empty()) arg = "<" + arg + ">";
    if (!def.empty()) def = "[" + def + "]";
    std::snprintf(buf, sizeof(buf), " -%c %-9s: %-54s%10s", opt, arg.c_str(), text.c_str(),
                  def.c_str());
    help_.push_back(buf);
  }

  void Handle(int opt, const char* a) {
    switch (opt) {
      case 'f': filename_ = a; break;
      case 'g': scale_ = std::atoi(a); break;
      case 'h': PrintUsage(); break;
      case 'k': degree_ = std::atoi(a); break;
      case 's': symmetrize_ = true; break;
      case 'u': uniform_ = true; scale_ = std::atoi(a); break;
      case 'm': in_place_ = true; break;
      case 'z': seed_ = std::atoll(a); break;
      case 'c': grid_side_ = std::atoll(a); break;
      case 'x': device_ = std::atoi(a); break;
      case 'a': do_analysis_ = true; break;
      case 'n': num_trials_ = std::atoi(a); break;
      case 'r': start_vertex_ = std::atol(a); break;
      case 'v': do_verify_ = true; break;
      case 'l': logging_ = true; break;
      case 'j': batch_ = std::max(1, std::atoi(a)); break;
      case 'p': precision_ = a; break;
      case 'q': autotune_ = true; break;
      case 'o': order_by_degree_ = true; break;
      case 'i':
        num_iters_ = std::atoi(a);
        max_iters_ = num_iters_;
        break;
      case 't': tolerance_ = std::stod(a); break;
      case 'd': delta_ = static_cast<int32_t>(std::atol(a)); break;
      case 'b': out_sg_ = true; out_filename_ = a; break;
      case 'e': out_el_ = true; out_filename_ = a; break;
      case 'w': out_weighted_ = true; break;
      default: break;
    }
  }

  int argc_;
  char** argv_;
  std::string name_;
  Kind kind_;
  std::string opts_;
  std::vector<std::string> help_;
  int scale_ = -1;
  int degree_ = 16;
  std::string filename_;
  bool symmetrize_ = false;
  bool uniform_ = false;
  bool in_place_ = false;
  int64_t seed_ = kRandSeed;
  int64_t grid_side_ = -1;
  int device_ = 0;
  bool do_analysis_ = false;
  int num_trials_ = 16;
  int64_t start_vertex_ = -1;
  bool do_verify_ = false;
  bool logging_ = false;
  int batch_ = 1;
  std::string precision_ = "fp32";
  bool autotune_ = false;
  bool order_by_degree_ = false;
  int num_iters_ = 1;
  int max_iters_ = 20;
  double tolerance_ = 1e-4;
  int32_t delta_ = 1;
  std::string out_filename_;
  bool out_weighted_ = false;
  bool out_el_ = false;
  bool out_sg_ = false;
};

// Edge list for the synthetic graph the flags ask for.
inline Buf<Edge> GenerateEdgeList(const CLI& cli, Buf<WeightT>* grid_weights) {
  Timer t;
  t.Start();
  Buf<Edge> el;
  if (cli.grid_side() != -1) {
    GridSpec sp;
    sp.side = cli.grid_side();
    sp.seed = cli.seed();
    el = MakeGridEL(sp, grid_weights);
  } else if (cli.uniform()) {
    el = MakeUniformEL(cli.scale(), cli.degree(), cli.seed());
  } else {
    el = MakeKroneckerEL(cli.scale(), cli.degree(), cli.seed());
  }
  t.Stop();
  PrintTime("Generate Time", t.Seconds());
  return el;
}

// Graph acquisition as the drivers see it: a serialized file, or a
// generated graph (always symmetrised).
template <class DestT>
CSRGraphT<DestT> MakeGraph(const CLI& cli) {
  constexpr bool kWeighted = std::is_same<DestT, WNode>::value;
  if (cli.in_place() && kWeighted)
    throw Error(-30, "In-place building (-m) does not support weighted graphs");
  if (!cli.filename().empty()) {
    const std::string suffix = FileSuffix(cli.filename());
    if (suffix != ".sg" && suffix != ".wsg") throw Error(-3, "Unrecognized suffix: " + suffix);
    Timer t;
    t.Start();
    CSRGraphT<DestT> g = ReadSerialized<DestT>(cli.filename());
    t.Stop();
    PrintTime("Read Time", t.Seconds());
    return g;
  }
  Buf<WeightT> grid_w;
  Buf<Edge> el = GenerateEdgeList(cli, &grid_w);
  Timer t;
  t.Start();
  Buf<WeightT> w;
  if (kWeighted) w = cli.grid_side() != -1 ? std::move(grid_w) : MakeWeights(el.size(), cli.seed());
  CSRGraphT<DestT> g = BuildFromEdgeList<DestT>(std::move(el), kWeighted ? w.data() : nullptr,
                                                cli.symmetrize());
  t.Stop();
  PrintTime("Build Time", t.Seconds());
  return g;
}

// Roots: uniformly drawn vertices with non-zero out-degree, from an
// mt19937_64 seeded with kRandSeed, or always the given vertex.
template <class GraphT>
class SourcePicker {
 public:
  explicit SourcePicker(const GraphT& g, NodeID given = -1)
      : g_(g), given_(given), rng_(kRandSeed), draw_(static_cast<int32_t>(g.num_nodes() - 1), rng_) {}

  NodeID PickNext() {
    if (given_ != -1) return given_;
    if (g_.num_edges_directed() == 0) throw Error(-41, "graph has no edges to pick roots from");
    NodeID s;
    do {
      s = draw_();
    } while (g_.out_degree(s) == 0);
    return s;
  }

 private:
  const GraphT& g_;
  NodeID given_;
  std::mt19937_64 rng_;
  BoundedDraw<std::mt19937_64> draw_;
};

// Relabel by degree only for dense, skewed graphs: average degree >= 10 and
// the sampled mean degree / 1.3 above the sampled median (1000 samples drawn
// with SourcePicker), as WorthRelabelling does (src/tc.cc:75-91).
inline bool WorthRelabelling(const Graph& g) {
  const int64_t avg = g.num_nodes() ? g.num_edges() / g.num_nodes() : 0;
  if (avg < 10) return false;
  SourcePicker<Graph> sp(g);
  const int64_t k = std::min<int64_t>(1000, g.num_nodes());
  std::vector<int64_t> s(static_cast<size_t>(k));
  int64_t total = 0;
  for (int64_t i = 0; i < k; i++) {
    s[i] = g.out_degree(sp.PickNext());
    total += s[i];
  }
  std::sort(s.begin(), s.end());
  const double mean = static_cast<double>(total) / k;
  const double median = static_cast<double>(s[k / 2]);
  return mean / 1.3 > median;
}

// The k largest (value, key) pairs, largest first.
template <class KeyT, class ValT>
std::vector<std::pair<ValT, KeyT>> TopK(const std::vector<std::pair<KeyT, ValT>>& kv, size_t k) {
  std::vector<std::pair<ValT, KeyT>> top;
  top.reserve(kv.size());
  for (const auto& p : kv) top.emplace_back(p.second, p.first);
  const size_t kk = std::min(k, top.size());
  std::partial_sort(top.begin(), top.begin() + kk, top.end(),
                    std::greater<std::pair<ValT, KeyT>>());
  top.resize(kk);
  return top;
}

// Runs and times the kernel num_trials times; analysis after the last trial,
// verification after every trial.
template <class GraphT, class Kernel, class Stats, class Verify>
void BenchmarkKernel(const CLI& cli, const GraphT& g, Kernel kernel, Stats stats, Verify verify) {
  g.PrintStats();
  double total = 0;
  Timer t;
  for (int iter = 0; iter < cli.num_trials(); iter++) {
    t.Start();
    auto result = kernel(g);
    t.Stop();
    PrintTime("Trial Time", t.Seconds());
    total += t.Seconds();
    if (cli.do_analysis() && iter == cli.num_trials() - 1) stats(g, result);
    if (cli.do_verify()) {
      t.Start();
      PrintLabel("Verification", verify(g, result) ? "PASS" : "FAIL");
      t.Stop();
      PrintTime("Verification Time", t.Seconds());
    }
  }
  PrintTime("Average Time", cli.num_trials() ? total / cli.num_trials() : 0.0);
}

}  // namespace gap
