// Synthetic edge-list generators.
//
// Reference: src/generator.h.  The Kronecker (R-MAT) and uniform generators
// must produce exactly the reference's edge lists so that a graph built here
// is the graph the reference builds from the same flags:
//   * bounded draws with the reference's rejection cutoff and 32-bit
//     truncation (UniDist, generator.h:34-64);
//   * edges generated in blocks of 2^18, each block's engine reseeded with
//     seed + block index, so the output does not depend on the thread count
//     (generator.h:105-150);
//   * R-MAT quadrant thresholds A=0.57, B=C=0.19 taken as uint32 fractions of
//     2^32-1 and the quadrant rule of generator.h:136-145;
//   * vertex IDs permuted by std::shuffle with an mt19937 seeded with the
//     seed (generator.h:93-103) -- the same libstdc++ routine is called here;
//   * weights uniform in [1, 255] from the same per-block streams
//     (generator.h:170-187).
// The reference hard-wires the seed to kRandSeed; here it is a parameter
// (default kRandSeed) so that the other benchmark graphs can be generated.
// The 2D grid and the packed-graph scale draw have no reference counterpart.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <random>
#include <sstream>
#include <vector>

#include "gap_core.h"

namespace gap {

struct Edge {
  NodeID u;
  NodeID v;
};

// Draws uniformly from [0, max_value] the way the reference's UniDist does,
// including its truncation of 64-bit engine output to 32 bits.
template <class Rng>
class BoundedDraw {
 public:
  BoundedDraw(int32_t max_value, Rng& rng) : rng_(rng) {
    using R = typename Rng::result_type;
    const R rmax = Rng::max();
    no_mod_ = rmax == static_cast<R>(static_cast<uint32_t>(max_value));
    mod_ = static_cast<uint32_t>(max_value) + 1u;
    const uint32_t rem = static_cast<uint32_t>(rmax % mod_);
    cutoff_ = (rem == mod_ - 1u) ? 0u : static_cast<uint32_t>(rmax - rem);
  }
  int32_t operator()() {
    uint32_t r = static_cast<uint32_t>(rng_());
    if (no_mod_) return static_cast<int32_t>(r);
    if (cutoff_ != 0)
      while (r >= cutoff_) r = static_cast<uint32_t>(rng_());
    return static_cast<int32_t>(r % mod_);
  }

 private:
  Rng& rng_;
  bool no_mod_;
  uint32_t mod_;
  uint32_t cutoff_;
};

constexpr int64_t kGenBlock = int64_t(1) << 18;

inline void CheckScale(int scale) {
  if (scale < 0 || scale > 30) {
    std::ostringstream os;
    os << "NodeID type (max: " << std::numeric_limits<NodeID>::max() << ") too small to hold "
       << (int64_t(1) << std::min(scale, 62));
    throw Error(-31, os.str());
  }
}

// Relabels every endpoint through a random permutation of [0, n).
inline void PermuteIDs(Buf<Edge>& el, int64_t n, int64_t seed) {
  Buf<NodeID> perm(static_cast<size_t>(n));
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++) perm[i] = static_cast<NodeID>(i);
  std::mt19937 rng(static_cast<std::mt19937::result_type>(seed));
  std::shuffle(perm.begin(), perm.end(), rng);
  const int64_t m = static_cast<int64_t>(el.size());
#pragma omp parallel for
  for (int64_t e = 0; e < m; e++) el[e] = Edge{perm[el[e].u], perm[el[e].v]};
}

// Graph500 R-MAT edge list of degree * 2^scale edges, IDs permuted.
inline Buf<Edge> MakeKroneckerEL(int scale, int degree, int64_t seed = kRandSeed) {
  CheckScale(scale);
  const int64_t n = int64_t(1) << scale;
  const int64_t m = n * degree;
  const uint32_t kMax = std::numeric_limits<uint32_t>::max();
  const uint32_t A = static_cast<uint32_t>(0.57 * kMax);
  const uint32_t B = static_cast<uint32_t>(0.19 * kMax);
  const uint32_t C = static_cast<uint32_t>(0.19 * kMax);
  const uint32_t AB = A + B, ABC = A + B + C;
  Buf<Edge> el(static_cast<size_t>(m));
#pragma omp parallel
  {
    std::mt19937 rng;
#pragma omp for schedule(static)
    for (int64_t blk = 0; blk < (m + kGenBlock - 1) / kGenBlock; blk++) {
      rng.seed(static_cast<std::mt19937::result_type>(seed + blk));
      const int64_t lo = blk * kGenBlock, hi = std::min(lo + kGenBlock, m);
      for (int64_t e = lo; e < hi; e++) {
        uint32_t s = 0, d = 0;
        for (int lvl = 0; lvl < scale; lvl++) {
          const uint32_t r = rng();
          s <<= 1;
          d <<= 1;
          if (r < AB) {
            d += (r > A);
          } else {
            s += 1;
            d += (r > ABC);
          }
        }
        el[e] = Edge{static_cast<NodeID>(s), static_cast<NodeID>(d)};
      }
    }
  }
  PermuteIDs(el, n, seed);
  return el;
}

// Uniform-random edge list of degree * 2^scale edges (no permutation).
inline Buf<Edge> MakeUniformEL(int scale, int degree, int64_t seed = kRandSeed) {
  CheckScale(scale);
  const int64_t n = int64_t(1) << scale;
  const int64_t m = n * degree;
  Buf<Edge> el(static_cast<size_t>(m));
#pragma omp parallel
  {
    std::mt19937 rng;
    BoundedDraw<std::mt19937> draw(static_cast<int32_t>(n - 1), rng);
#pragma omp for schedule(static)
    for (int64_t blk = 0; blk < (m + kGenBlock - 1) / kGenBlock; blk++) {
      rng.seed(static_cast<std::mt19937::result_type>(seed + blk));
      const int64_t lo = blk * kGenBlock, hi = std::min(lo + kGenBlock, m);
      for (int64_t e = lo; e < hi; e++) {
        // The reference builds Edge(udist(), udist()); g++ evaluates the
        // two constructor arguments right to left, so the first draw is the
        // destination.
        const NodeID v = draw();
        const NodeID u = draw();
        el[e] = Edge{u, v};
      }
    }
  }
  return el;
}

// One weight per edge-list entry, uniform in [1, max_weight].
inline Buf<WeightT> MakeWeights(int64_t m, int64_t seed = kRandSeed, int32_t max_weight = 255) {
  Buf<WeightT> w(static_cast<size_t>(m));
#pragma omp parallel
  {
    std::mt19937 rng;
    BoundedDraw<std::mt19937> draw(max_weight - 1, rng);
#pragma omp for schedule(static)
    for (int64_t blk = 0; blk < (m + kGenBlock - 1) / kGenBlock; blk++) {
      rng.seed(static_cast<std::mt19937::result_type>(seed + blk));
      const int64_t lo = blk * kGenBlock, hi = std::min(lo + kGenBlock, m);
      for (int64_t e = lo; e < hi; e++) w[e] = static_cast<WeightT>(draw() + 1);
    }
  }
  return w;
}

// Counter-based 64-bit mixer (SplitMix64 finaliser) for the generators
// below, which are parallel and need no sequential stream.
inline uint64_t Mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// High-diameter test graph: side x side 4-neighbour lattice (vertex r*side+c),
// each lattice edge deleted with probability delete_frac, plus
// round(shortcut_frac * lattice edges) uniformly random extra pairs.
// Weights (if requested) are uniform in [1, max_weight].
struct GridSpec {
  int64_t side = 4096;
  double delete_frac = 0.05;
  double shortcut_frac = 0.01;
  int64_t seed = 7;
  int32_t max_weight = 1000;
};

inline Buf<Edge> MakeGridEL(const GridSpec& sp, Buf<WeightT>* weights) {
  const int64_t L = sp.side;
  if (L < 2 || L * L > std::numeric_limits<NodeID>::max()) throw Error(-31, "grid side out of range");
  const int64_t lattice = 2 * L * (L - 1);
  const int64_t shortcuts = std::llround(sp.shortcut_frac * static_cast<double>(lattice));
  // Pass 1: which lattice edges survive.
  Buf<uint8_t> keep(static_cast<size_t>(lattice));
  const uint64_t s0 = Mix64(static_cast<uint64_t>(sp.seed) ^ 0x6772696400000000ull);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < lattice; i++) {
    const double u01 = static_cast<double>(Mix64(s0 + 2 * static_cast<uint64_t>(i)) >> 11) * 0x1.0p-53;
    keep[i] = u01 >= sp.delete_frac;
  }
  Buf<int64_t> cnt(static_cast<size_t>(lattice));
  int64_t kept = 0;
  for (int64_t i = 0; i < lattice; i++) {
    cnt[i] = kept;
    kept += keep[i];
  }
  const int64_t m = kept + shortcuts;
  Buf<Edge> el(static_cast<size_t>(m));
  const int64_t horiz = L * (L - 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < lattice; i++) {
    if (!keep[i]) continue;
    int64_t u, v;
    if (i < horiz) {  // (r, c) - (r, c+1)
      const int64_t r = i / (L - 1), c = i % (L - 1);
      u = r * L + c;
      v = u + 1;
    } else {  // (r, c) - (r+1, c)
      const int64_t j = i - horiz;
      const int64_t r = j / L, c = j % L;
      u = r * L + c;
      v = u + L;
    }
    el[cnt[i]] = Edge{static_cast<NodeID>(u), static_cast<NodeID>(v)};
  }
  const uint64_t s1 = Mix64(static_cast<uint64_t>(sp.seed) ^ 0x73686f7274637574ull);
  const uint64_t nv = static_cast<uint64_t>(L * L);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < shortcuts; k++) {
    const uint64_t a = Mix64(s1 + 2 * static_cast<uint64_t>(k));
    const uint64_t b = Mix64(s1 + 2 * static_cast<uint64_t>(k) + 1);
    el[kept + k] = Edge{static_cast<NodeID>(a % nv), static_cast<NodeID>(b % nv)};
  }
  if (weights) {
    weights->reset(static_cast<size_t>(m));
    const uint64_t s2 = Mix64(static_cast<uint64_t>(sp.seed) ^ 0x7765696768747321ull);
    const uint64_t range = static_cast<uint64_t>(sp.max_weight);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; e++)
      (*weights)[e] = static_cast<WeightT>(1 + Mix64(s2 + static_cast<uint64_t>(e)) % range);
  }
  return el;
}

// Scale of packed graph i in the packed small-graph workload: uniform in
// [lo, hi] from an engine seeded with i.
inline int PackedGraphScale(int64_t i, int lo = 10, int hi = 16) {
  std::mt19937_64 rng(static_cast<uint64_t>(i));
  return lo + static_cast<int>(rng() % static_cast<uint64_t>(hi - lo + 1));
}

}  // namespace gap
