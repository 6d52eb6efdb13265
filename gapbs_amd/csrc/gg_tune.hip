// Per-graph autotuner (gg_autotune, gg_tuning_get, gg_tuning_set).
//
// No reference counterpart: the reference exposes alpha/beta as DOBFS
// defaults (src/bfs.cc:122-123) and delta as sssp's -d flag (src/sssp.cc:
// 87-88, command_line.h); everything else is fixed.  Here each candidate
// setting is run on the device graph from the caller's roots and timed with
// the handle's own HIP-event device time; the fastest is kept in the handle.
// Results do not depend on any of these settings (every kernel's parity
// holds for all of them), only the time does.
#include <algorithm>
#include <iterator>

#include "gg_common.h"

namespace gg {
namespace {

constexpr int kAlphas[] = {2, 4, 8, 15, 30, 60, 120, 240};
constexpr int kBetas[] = {2, 6, 18, 54, 162};
constexpr int kBatchAlphas[] = {1, 2, 4, 8, 15, 30, 60, 120, 1000};
constexpr int kSmallMaxes[] = {8, 16, 32, 64, 128};
constexpr int kMediumMaxes[] = {256, 512, 1024, 2048, 8192};
constexpr int kWidths[] = {4, 8, 16, 32};
constexpr int kPrIters = 3;

void Check(int rc) {
  if (rc != GG_OK) throw Failure(rc, gg_last_error());
}

// Device ms of one call; the best of two runs damps launch-queue noise.
template <class F>
double Time(F&& run) {
  double best = 1e300;
  for (int rep = 0; rep < 2; rep++) {
    gg_stats st;
    Check(run(&st));
    best = std::min(best, st.device_ms);
  }
  return best;
}

void TuneBfs(gg_graph* g, const int32_t* roots, int k) {
  const int kk = std::min(k, 4);
  double best = 1e300;
  for (int a : kAlphas)
    for (int b : kBetas) {
      double t = 0;
      for (int i = 0; i < kk; i++)
        t += Time([&](gg_stats* st) { return gg_bfs(g, roots[i], a, b, nullptr, nullptr, st); });
      if (t < best) {
        best = t;
        g->tune.bfs_alpha = a;
        g->tune.bfs_beta = b;
      }
    }
  g->tune.best_ms[0] = best / kk;
  g->tune.tuned |= GG_TUNE_BFS;
}

void TuneBatch(gg_graph* g, const int32_t* roots, int k) {
  const int kk = std::min(k, 64);
  double best = 1e300;
  for (int a : kBatchAlphas) {
    const double t = Time([&](gg_stats* st) { return gg_bfs_batch(g, roots, kk, a, nullptr, st); });
    if (t < best) {
      best = t;
      g->tune.batch_alpha = a;
    }
  }
  // On a graph with propagation blocks and the default block policy: the
  // first bottom-up level with and without the blocks pass, at that alpha.
  int64_t info[6];
  Check(gg_pagerank_blocks(g, GG_PR_FP32, info));
  if (info[0] > 0 && g->tune.bfs_blocks_pct == 0) {
    int choice = 0;
    for (int pct : {50, -1}) {
      g->tune.bfs_blocks_pct = pct;
      const double t = Time([&](gg_stats* st) { return gg_bfs_batch(g, roots, kk, g->tune.batch_alpha, nullptr, st); });
      if (t < best) {
        best = t;
        choice = pct;
      }
    }
    g->tune.bfs_blocks_pct = choice;
  }
  g->tune.best_ms[1] = best;
  g->tune.tuned |= GG_TUNE_BATCH;
}

// Candidates from 256 up to 4096, then down to 1.  Going down, a width is
// skipped once the time the last measured one projects for it (the bucket
// count grows like 1 / delta, and small widths are bound by per-bucket
// overhead) exceeds 3x the best: on high-diameter graphs delta = 1 would
// otherwise take minutes.
void TuneSssp(gg_graph* g, const int32_t* roots, int k) {
  GG_REQUIRE(g->weighted, GG_ENOTSUP, "SSSP tuning needs a weighted upload");
  const int kk = std::min(k, 2);
  double best = 1e300;
  auto measure = [&](int delta) {
    double t = 0;
    for (int i = 0; i < kk && t < 3 * best; i++) {
      gg_stats st;
      Check(gg_sssp(g, roots[i], delta, nullptr, &st));
      t += st.device_ms;
    }
    if (t < best) {
      best = t;
      g->tune.sssp_delta = delta;
    }
    return t;
  };
  for (int delta = 256; delta <= 4096; delta *= 2) measure(delta);
  double t_last = 0;
  int d_last = 0;
  for (int delta = 128; delta >= 1; delta /= 2) {
    if (d_last && t_last * (static_cast<double>(d_last) / delta) > 3 * best) break;
    t_last = measure(delta);
    d_last = delta;
  }
  g->tune.best_ms[2] = best / kk;
  g->tune.tuned |= GG_TUNE_SSSP;
}

// Graphs whose Jacobi runs go through propagation blocks at the default
// width: the width and its two neighbours (at most 8192 tiles) are timed
// (device time excludes each width's layout build) and the fastest kept.
// Returns false when the graph runs without blocks.
bool TunePrBlocks(gg_graph* g) {
  int64_t info[6];
  Check(gg_pagerank_blocks(g, GG_PR_FP32, info));
  if (info[0] == 0 || g->tune.pr_blocks_log2 != 0) return false;
  const int d = static_cast<int>(info[0]);
  auto run = [&]() { return Time([&](gg_stats* st) { return gg_pagerank(g, kPrIters, 0.0, GG_PR_FP32, nullptr, st); }); };
  double best = run();
  int choice = 0;
  for (int lg : {d - 1, d + 1}) {
    if (lg < 4 || lg > 14 || ((g->n + (int64_t(1) << lg) - 1) >> lg) > 8192) continue;
    g->tune.pr_blocks_log2 = lg;
    const double t = run();
    if (t < best) {
      best = t;
      choice = lg;
    }
  }
  g->tune.pr_blocks_log2 = choice;  // 0: the default width
  g->tune.best_ms[3] = best / kPrIters;
  g->tune.tuned |= GG_TUNE_PR;
  return true;
}

void TunePr(gg_graph* g) {
  if (TunePrBlocks(g)) return;
  auto run = [&]() {
    g->pr_bins[0] = -1;  // the cut-offs changed: rebuild the bins
    return Time([&](gg_stats* st) { return gg_pagerank(g, kPrIters, 0.0, GG_PR_FP32, nullptr, st); });
  };
  double best = run();
  auto descend = [&](int32_t* knob, const int* values, int count) {
    const int32_t keep = *knob;
    int32_t choice = keep;
    for (int i = 0; i < count; i++) {
      if (values[i] == keep) continue;
      *knob = values[i];
      if (g->tune.pr_small_max >= g->tune.pr_medium_max) continue;
      const double t = run();
      if (t < best) {
        best = t;
        choice = values[i];
      }
    }
    *knob = choice;
  };
  descend(&g->tune.pr_small_max, kSmallMaxes, static_cast<int>(std::size(kSmallMaxes)));
  descend(&g->tune.pr_medium_max, kMediumMaxes, static_cast<int>(std::size(kMediumMaxes)));
  descend(&g->tune.pr_blocks_per_cu, kWidths, static_cast<int>(std::size(kWidths)));
  g->pr_bins[0] = -1;
  g->tune.best_ms[3] = best / kPrIters;
  g->tune.tuned |= GG_TUNE_PR;
}

}  // namespace
}  // namespace gg

using namespace gg;

extern "C" int gg_autotune(gg_graph* g, int what, const int32_t* roots, int k, gg_tuning* out) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr, GG_EINVAL, "null graph");
    GG_REQUIRE((what & ~(GG_TUNE_BFS | GG_TUNE_BATCH | GG_TUNE_SSSP | GG_TUNE_PR)) == 0, GG_EINVAL,
               "unknown tuning selection");
    const bool need_roots = what & (GG_TUNE_BFS | GG_TUNE_BATCH | GG_TUNE_SSSP);
    GG_REQUIRE(!need_roots || (k >= 1 && roots), GG_EINVAL, "tuning needs at least one root");
    for (int i = 0; i < k; i++) GG_REQUIRE(roots[i] >= 0 && roots[i] < g->n, GG_EINVAL, "root out of range");
    const bool prof = g->profile;
    g->profile = false;  // tuning runs are not part of any measurement
    const gg_tuning saved = g->tune;
    try {
      if (what & GG_TUNE_BFS) TuneBfs(g, roots, k);
      if (what & GG_TUNE_BATCH) TuneBatch(g, roots, k);
      if (what & GG_TUNE_SSSP) TuneSssp(g, roots, k);
      if (what & GG_TUNE_PR) TunePr(g);
    } catch (...) {
      g->tune = saved;
      g->pr_bins[0] = -1;
      g->profile = prof;
      throw;
    }
    g->profile = prof;
    if (out) *out = g->tune;
  });
}

extern "C" int gg_tuning_get(const gg_graph* g, gg_tuning* out) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr && out != nullptr, GG_EINVAL, "null argument");
    *out = g->tune;
  });
}

extern "C" int gg_tuning_set(gg_graph* g, const gg_tuning* in) {
  return Guard([&] {
    GG_REQUIRE(g != nullptr && in != nullptr, GG_EINVAL, "null argument");
    GG_REQUIRE(in->bfs_alpha > 0 && in->bfs_beta > 0 && in->batch_alpha > 0 && in->sssp_delta > 0, GG_EINVAL,
               "alpha, beta and delta must be positive");
    GG_REQUIRE(in->pr_small_max >= 1 && in->pr_small_max < in->pr_medium_max, GG_EINVAL,
               "PageRank cut-offs must satisfy 1 <= small < medium");
    GG_REQUIRE(in->pr_blocks_per_cu >= 1 && in->pr_blocks_per_cu <= 64, GG_EINVAL, "launch width out of range");
    for (int i = 0; i < 2; i++) {
      GG_REQUIRE(in->pr_segment_log2[i] == -1 || in->pr_segment_log2[i] == 0 ||
                     (in->pr_segment_log2[i] >= 4 && in->pr_segment_log2[i] <= 30),
                 GG_EINVAL, "PageRank segment width out of range");
      GG_REQUIRE(in->pr_lds_log2[i] == -1 || in->pr_lds_log2[i] == 0 ||
                     (in->pr_lds_log2[i] >= 4 && in->pr_lds_log2[i] <= (i == 0 ? 15 : 14)),
                 GG_EINVAL, "PageRank LDS segment width out of range");
    }
    GG_REQUIRE(in->pr_seg_min_density >= 0 && in->pr_seg_min_density <= 64, GG_EINVAL,
               "PageRank segment density out of range");
    GG_REQUIRE(in->bfs_blocks_pct >= -1 && in->bfs_blocks_pct <= 100, GG_EINVAL,
               "batched BFS block threshold out of range");
    GG_REQUIRE(in->pr_blocks_log2 == -1 || in->pr_blocks_log2 == 0 ||
                   (in->pr_blocks_log2 >= 4 && in->pr_blocks_log2 <= 14),
               GG_EINVAL, "PageRank block width out of range");
    const bool bins_changed =
        in->pr_small_max != g->tune.pr_small_max || in->pr_medium_max != g->tune.pr_medium_max;
    g->tune = *in;
    if (bins_changed) g->pr_bins[0] = -1;
  });
}
