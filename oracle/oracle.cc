// CPU oracle for the kernel hot path -- TEST INFRASTRUCTURE ONLY.
// See oracle.h.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline load this library; the product never does.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstring>
#include <functional>
#include <limits>
#include <queue>
#include <random>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

inline double Since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

inline bool Bit(const std::vector<uint64_t>& bm, int64_t i) { return (bm[i >> 6] >> (i & 63)) & 1; }

}  // namespace

extern "C" {

// src/bfs.cc:204-215: plain BFS from src over out-neighbours.
void orc_bfs_depth(int64_t n, const int64_t* off, const int32_t* dst, int32_t src, int32_t* depth) {
  std::fill(depth, depth + n, -1);
  std::vector<int32_t> order;
  order.reserve(static_cast<size_t>(n));
  depth[src] = 0;
  order.push_back(src);
  for (size_t i = 0; i < order.size(); i++) {
    const int32_t u = order[i];
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = dst[e];
      if (depth[v] == -1) {
        depth[v] = depth[u] + 1;
        order.push_back(v);
      }
    }
  }
}

// Restates DOBFS (src/bfs.cc:122-179) with its steps TDStep (67-89),
// BUStep (46-64), QueueToBitmap/BitmapToQueue (92-112) and InitParent
// (114-120).  Frontier appends go through per-thread buffers; bottom-up
// works on whole 64-vertex bitmap words so `next` needs no atomics.
double orc_dobfs(int64_t n, const int64_t* off, const int32_t* dst, const int64_t* ioff,
                 const int32_t* idst, int64_t e_dir, int32_t src, int alpha, int beta,
                 int32_t* parent, orc_bfs_stats* st) {
  return orc_dobfs_trace(n, off, dst, ioff, idst, e_dir, src, alpha, beta, parent, st, nullptr, 0,
                         nullptr);
}

double orc_dobfs_trace(int64_t n, const int64_t* off, const int32_t* dst, const int64_t* ioff,
                       const int32_t* idst, int64_t e_dir, int32_t src, int alpha, int beta,
                       int32_t* parent, orc_bfs_stats* st, orc_bfs_step* steps, int max_steps,
                       int* nsteps) {
  const auto t0 = Clock::now();
  int ns = 0;
  auto record = [&](int bu, int64_t fr, int64_t vx, int64_t ex, int64_t disc) {
    if (steps && ns < max_steps) steps[ns] = orc_bfs_step{bu, 0, fr, vx, ex, disc};
    ns++;
  };
  orc_bfs_stats s;
  std::memset(&s, 0, sizeof(s));
  const int64_t nw = (n + 63) / 64;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const int64_t d = off[i + 1] - off[i];
    parent[i] = d != 0 ? static_cast<int32_t>(-d) : -1;
  }
  parent[src] = src;
  std::vector<int32_t> queue(static_cast<size_t>(n));
  size_t qb = 0, qe = 1;
  queue[0] = src;
  std::vector<uint64_t> front(static_cast<size_t>(nw), 0), next(static_cast<size_t>(nw), 0);
  int64_t edges_to_check = e_dir;
  int64_t scout = off[src + 1] - off[src];
  s.reached = 1;
  s.reached_edges = scout;
  while (qe > qb) {
    if (scout > edges_to_check / alpha) {
      std::fill(front.begin(), front.end(), 0);
      for (size_t i = qb; i < qe; i++) front[queue[i] >> 6] |= uint64_t(1) << (queue[i] & 63);
      int64_t awake = static_cast<int64_t>(qe - qb), old_awake;
      do {
        old_awake = awake;
        int64_t aw = 0, ex = 0, vx = 0, re = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : aw, ex, vx, re)
        for (int64_t w = 0; w < nw; w++) {
          uint64_t word = 0;
          const int64_t hi = std::min<int64_t>(64, n - w * 64);
          for (int64_t l = 0; l < hi; l++) {
            const int64_t u = w * 64 + l;
            if (parent[u] >= 0) continue;
            vx++;
            for (int64_t e = ioff[u]; e < ioff[u + 1]; e++) {
              ex++;
              const int32_t v = idst[e];
              if (Bit(front, v)) {
                parent[u] = v;
                word |= uint64_t(1) << l;
                aw++;
                re += off[u + 1] - off[u];
                break;
              }
            }
          }
          next[w] = word;
        }
        record(1, old_awake, vx, ex, aw);
        awake = aw;
        s.edges_examined += ex;
        s.vertices_examined += vx;
        s.reached += aw;
        s.reached_edges += re;
        s.bu_steps++;
        front.swap(next);
      } while (awake >= old_awake || awake > n / beta);
      qb = qe = 0;
      for (int64_t w = 0; w < nw; w++) {
        uint64_t word = front[w];
        while (word) {
          const int l = __builtin_ctzll(word);
          queue[qe++] = static_cast<int32_t>(w * 64 + l);
          word &= word - 1;
        }
      }
      scout = 1;
    } else {
      edges_to_check -= scout;
      int64_t sc = 0, ex = 0, re = 0, found = 0;
      size_t tail = qe;
#pragma omp parallel reduction(+ : sc, ex, re, found)
      {
        std::vector<int32_t> local;
#pragma omp for schedule(dynamic, 64) nowait
        for (size_t i = qb; i < qe; i++) {
          const int32_t u = queue[i];
          ex += off[u + 1] - off[u];
          for (int64_t e = off[u]; e < off[u + 1]; e++) {
            const int32_t v = dst[e];
            int32_t cur = __atomic_load_n(&parent[v], __ATOMIC_RELAXED);
            if (cur < 0 && __atomic_compare_exchange_n(&parent[v], &cur, u, false, __ATOMIC_RELAXED,
                                                       __ATOMIC_RELAXED)) {
              local.push_back(v);
              sc += -static_cast<int64_t>(cur);
              re += off[v + 1] - off[v];
              found++;
            }
          }
        }
        size_t at;
#pragma omp atomic capture
        {
          at = tail;
          tail += local.size();
        }
        std::copy(local.begin(), local.end(), queue.begin() + at);
      }
      record(0, static_cast<int64_t>(qe - qb), static_cast<int64_t>(qe - qb), ex, found);
      s.vertices_examined += static_cast<int64_t>(qe - qb);
      s.edges_examined += ex;
      s.reached += found;
      s.reached_edges += re;
      s.td_steps++;
      scout = sc;
      qb = qe;
      qe = tail;
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    if (parent[i] < -1) parent[i] = -1;
  if (st) *st = s;
  if (nsteps) *nsteps = ns;
  return Since(t0);
}

void orc_dijkstra(int64_t n, const int64_t* off, const int32_t* wdst, int32_t src, int32_t* dist) {
  const int32_t inf = INT32_MAX / 2;
  std::fill(dist, dist + n, inf);
  typedef std::pair<int32_t, int32_t> Item;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  dist[src] = 0;
  pq.push(Item(0, src));
  while (!pq.empty()) {
    const Item top = pq.top();
    pq.pop();
    const int32_t d = top.first, u = top.second;
    if (d != dist[u]) continue;
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = wdst[2 * e], w = wdst[2 * e + 1];
      if (d + w < dist[v]) {
        dist[v] = d + w;
        pq.push(Item(d + w, v));
      }
    }
  }
}

// Shortest-path certificate over a symmetric graph with positive weights
// (unit weights when wdst is null; dst then holds plain ids).  dist is the
// distance vector from src iff
//   (1) dist[src] == 0;
//   (2) every edge (u, v, w) out of a reached u has v reached and
//       dist[v] <= dist[u] + w (nothing relaxes, so dist <= the distances);
//   (3) every reached v != src has an edge to a reached u with
//       dist[u] + w == dist[v] (following such edges strictly lowers dist and
//       ends at src, so dist >= the distances).
// In a symmetric graph the edges out of v are the edges into v.  Returns the
// number of vertices breaking a condition (0: certified).
int64_t orc_distance_certificate(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* wdst,
                                 int32_t src, const int32_t* dist, int32_t unreached) {
  int64_t bad = (src < 0 || src >= n || dist[src] != 0) ? 1 : 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad)
  for (int64_t u = 0; u < n; u++) {
    const int64_t du = dist[u];
    const bool reached = du != unreached;
    if (reached && (du < 0 || (unreached >= 0 && du >= unreached))) {
      bad++;
      continue;
    }
    bool relaxes = false, tight = (u == src);
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = wdst ? wdst[2 * e] : dst[e];
      const int64_t w = wdst ? wdst[2 * e + 1] : 1;
      const int64_t dv = dist[v];
      if (w <= 0) relaxes = true;
      if (!reached) {
        if (dv != unreached) relaxes = true;  // a reached neighbour would reach u
        continue;
      }
      if (dv == unreached || dv > du + w) relaxes = true;
      if (dv != unreached && dv + w == du) tight = true;
    }
    if (relaxes || (reached && !tight)) bad++;
  }
  return bad;
}

// Restates DeltaStep (src/sssp.cc:87-153) and RelaxEdges (68-85).
double orc_delta_step(int64_t n, const int64_t* off, const int32_t* wdst, int64_t e_dir, int32_t src,
                      int32_t delta, int32_t* dist) {
  const auto t0 = Clock::now();
  const int32_t inf = INT32_MAX / 2;
  const size_t kMaxBin = std::numeric_limits<size_t>::max() / 2;
  const size_t kFuse = 1000;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) dist[i] = inf;
  dist[src] = 0;
  std::vector<int32_t> frontier(static_cast<size_t>(e_dir > 0 ? e_dir : 1));
  size_t shared_idx[2] = {0, kMaxBin};
  size_t tails[2] = {1, 0};
  frontier[0] = src;
  auto relax = [&](int32_t u, std::vector<std::vector<int32_t>>& bins) {
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = wdst[2 * e], w = wdst[2 * e + 1];
      int32_t old_d = __atomic_load_n(&dist[v], __ATOMIC_RELAXED);
      const int32_t new_d = __atomic_load_n(&dist[u], __ATOMIC_RELAXED) + w;
      while (new_d < old_d) {
        if (__atomic_compare_exchange_n(&dist[v], &old_d, new_d, false, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED)) {
          const size_t b = static_cast<size_t>(new_d / delta);
          if (b >= bins.size()) bins.resize(b + 1);
          bins[b].push_back(v);
          break;
        }
      }
    }
  };
#pragma omp parallel
  {
    std::vector<std::vector<int32_t>> bins;
    size_t iter = 0;
    while (shared_idx[iter & 1] != kMaxBin) {
      size_t& cur = shared_idx[iter & 1];
      size_t& nxt = shared_idx[(iter + 1) & 1];
      size_t& cur_tail = tails[iter & 1];
      size_t& nxt_tail = tails[(iter + 1) & 1];
#pragma omp for nowait schedule(dynamic, 64)
      for (size_t i = 0; i < cur_tail; i++) {
        const int32_t u = frontier[i];
        if (dist[u] >= delta * static_cast<int32_t>(cur)) relax(u, bins);
      }
      while (cur < bins.size() && !bins[cur].empty() && bins[cur].size() < kFuse) {
        std::vector<int32_t> copy;
        copy.swap(bins[cur]);
        for (int32_t u : copy) relax(u, bins);
      }
      for (size_t b = cur; b < bins.size(); b++)
        if (!bins[b].empty()) {
#pragma omp critical
          nxt = std::min(nxt, b);
          break;
        }
#pragma omp barrier
#pragma omp single nowait
      {
        cur = kMaxBin;
        cur_tail = 0;
      }
      if (nxt < bins.size()) {
        size_t at;
#pragma omp atomic capture
        {
          at = nxt_tail;
          nxt_tail += bins[nxt].size();
        }
        std::copy(bins[nxt].begin(), bins[nxt].end(), frontier.begin() + at);
        bins[nxt].clear();
      }
      iter++;
#pragma omp barrier
    }
  }
  return Since(t0);
}

}  // extern "C"

namespace {
template <class T>
int PageRankPull(int64_t n, const int64_t* in_off, const int32_t* in_dst, const int64_t* out_off,
                 int max_iters, double epsilon, T* scores, double* error_out) {
  // Constants computed in float exactly as pr_spmv.cc:32, 38-39 does.
  const float kDamp = 0.85f;
  const float init_f = 1.0f / n;
  const float base_f = (1.0f - kDamp) / n;
  const T init = init_f, base = base_f, damp = kDamp;
  std::vector<T> contrib(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) scores[i] = init;
  int iters = 0;
  double error = 0;
  for (int it = 0; it < max_iters; it++) {
    error = 0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) contrib[i] = scores[i] / static_cast<T>(out_off[i + 1] - out_off[i]);
#pragma omp parallel for reduction(+ : error) schedule(dynamic, 16384)
    for (int64_t u = 0; u < n; u++) {
      T total = 0;
      for (int64_t e = in_off[u]; e < in_off[u + 1]; e++) total += contrib[in_dst[e]];
      const T old = scores[u];
      // A product then a sum, as the reference's expression rounds.
      volatile T prod = damp * total;
      scores[u] = base + prod;
      error += std::fabs(static_cast<double>(scores[u]) - static_cast<double>(old));
    }
    iters++;
    if (error < epsilon) break;
  }
  *error_out = error;
  return iters;
}
}  // namespace

extern "C" {

int orc_pr_pull_f32(int64_t n, const int64_t* in_off, const int32_t* in_dst, const int64_t* out_off,
                    int max_iters, double epsilon, float* scores, double* error) {
  return PageRankPull<float>(n, in_off, in_dst, out_off, max_iters, epsilon, scores, error);
}

int orc_pr_pull_f64(int64_t n, const int64_t* in_off, const int32_t* in_dst, const int64_t* out_off,
                    int max_iters, double epsilon, double* scores, double* error) {
  return PageRankPull<double>(n, in_off, in_dst, out_off, max_iters, epsilon, scores, error);
}

namespace {
// cc.cc:41-55
void AfLink(int32_t u, int32_t v, int32_t* comp) {
  int32_t p1 = __atomic_load_n(&comp[u], __ATOMIC_RELAXED);
  int32_t p2 = __atomic_load_n(&comp[v], __ATOMIC_RELAXED);
  while (p1 != p2) {
    const int32_t high = p1 > p2 ? p1 : p2;
    const int32_t low = p1 + (p2 - high);
    int32_t p_high = __atomic_load_n(&comp[high], __ATOMIC_RELAXED);
    if (p_high == low) break;
    if (p_high == high &&
        __atomic_compare_exchange_n(&comp[high], &p_high, low, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      break;
    p1 = __atomic_load_n(&comp[__atomic_load_n(&comp[high], __ATOMIC_RELAXED)], __ATOMIC_RELAXED);
    p2 = __atomic_load_n(&comp[low], __ATOMIC_RELAXED);
  }
}

// cc.cc:59-66
void AfCompress(int64_t n, int32_t* comp) {
#pragma omp parallel for schedule(dynamic, 16384)
  for (int64_t i = 0; i < n; i++)
    while (comp[i] != comp[comp[i]]) comp[i] = comp[comp[i]];
}

// cc.cc:69-92
int32_t AfSample(int64_t n, const int32_t* comp) {
  std::unordered_map<int32_t, int> counts(32);
  std::mt19937 gen;
  std::uniform_int_distribution<int32_t> dist(0, static_cast<int32_t>(n - 1));
  for (int i = 0; i < 1024; i++) counts[comp[dist(gen)]]++;
  auto best = std::max_element(counts.begin(), counts.end(),
                               [](const std::pair<const int32_t, int>& a,
                                  const std::pair<const int32_t, int>& b) { return a.second < b.second; });
  return best->first;
}
}  // namespace

double orc_afforest(int64_t n, const int64_t* off, const int32_t* dst, int directed,
                    const int64_t* ioff, const int32_t* idst, int neighbor_rounds, int32_t* comp) {
  const auto t0 = Clock::now();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) comp[i] = static_cast<int32_t>(i);
  for (int r = 0; r < neighbor_rounds; r++) {
#pragma omp parallel for schedule(dynamic, 16384)
    for (int64_t u = 0; u < n; u++)
      if (off[u + 1] - off[u] > r) AfLink(static_cast<int32_t>(u), dst[off[u] + r], comp);
    AfCompress(n, comp);
  }
  const int32_t c = AfSample(n, comp);
#pragma omp parallel for schedule(dynamic, 16384)
  for (int64_t u = 0; u < n; u++) {
    if (comp[u] == c) continue;
    const int64_t b = std::min(off[u] + neighbor_rounds, off[u + 1]);
    for (int64_t e = b; e < off[u + 1]; e++) AfLink(static_cast<int32_t>(u), dst[e], comp);
    if (directed)
      for (int64_t e = ioff[u]; e < ioff[u + 1]; e++) AfLink(static_cast<int32_t>(u), idst[e], comp);
  }
  AfCompress(n, comp);
  return Since(t0);
}

void orc_cc_min_labels(int64_t n, const int64_t* off, const int32_t* dst, int32_t* comp) {
  std::vector<int32_t> parent(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; i++) parent[i] = static_cast<int32_t>(i);
  std::function<int32_t(int32_t)> find = [&](int32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (int64_t u = 0; u < n; u++)
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t a = find(static_cast<int32_t>(u)), b = find(dst[e]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  for (int64_t i = 0; i < n; i++) comp[i] = find(static_cast<int32_t>(i));
}

}  // extern "C"

namespace {
// One root of Brandes: depth-first order of discovery, path counts in FP64,
// then dependencies farthest first into delta (zero where unreached).
template <class T>
void BrandesRoot(int64_t n, const int64_t* off, const int32_t* dst, int32_t src, std::vector<int32_t>& depth,
                 std::vector<double>& sigma, std::vector<int32_t>& order, T* delta) {
  std::fill(depth.begin(), depth.end(), -1);
  std::fill(sigma.begin(), sigma.end(), 0.0);
  std::fill(delta, delta + n, T(0));
  order.clear();
  depth[src] = 0;
  sigma[src] = 1;
  order.push_back(src);
  for (size_t i = 0; i < order.size(); i++) {
    const int32_t u = order[i];
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = dst[e];
      if (depth[v] == -1) {
        depth[v] = depth[u] + 1;
        order.push_back(v);
      }
      if (depth[v] == depth[u] + 1) sigma[v] += sigma[u];
    }
  }
  // Farthest first (reverse BFS order visits depths in decreasing order).
  for (size_t i = order.size(); i-- > 0;) {
    const int32_t u = order[i];
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = dst[e];
      if (depth[v] == depth[u] + 1) {
        volatile double term = (sigma[u] / sigma[v]) * (1 + delta[v]);
        delta[u] += term;
      }
    }
  }
}

// Roots run in parallel, one per thread, each into its own delta row; the
// rows are then added into scores in root order, so the sums are those of
// the serial loop over roots (adding the zero of an unreached vertex leaves
// a score unchanged).
template <class T>
void Brandes(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* roots, int k, T* scores) {
  std::fill(scores, scores + n, T(0));
  const int width = std::max(1, std::min(k, omp_get_max_threads()));
  std::vector<T> rows(static_cast<size_t>(width) * static_cast<size_t>(n));
  for (int r0 = 0; r0 < k; r0 += width) {
    const int batch = std::min(width, k - r0);
#pragma omp parallel num_threads(batch)
    {
      std::vector<int32_t> depth(static_cast<size_t>(n)), order;
      std::vector<double> sigma(static_cast<size_t>(n));
      order.reserve(static_cast<size_t>(n));
#pragma omp for schedule(static, 1)
      for (int j = 0; j < batch; j++)
        BrandesRoot(n, off, dst, roots[r0 + j], depth, sigma, order, rows.data() + static_cast<size_t>(j) * n);
    }
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; v++)
      for (int j = 0; j < batch; j++) scores[v] += rows[static_cast<size_t>(j) * n + v];
  }
  T biggest = 0;
  for (int64_t i = 0; i < n; i++) biggest = std::max(biggest, scores[i]);
  for (int64_t i = 0; i < n; i++) scores[i] = scores[i] / biggest;
}
}  // namespace

extern "C" {

void orc_brandes_f32(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* roots, int k,
                     float* scores) {
  Brandes<float>(n, off, dst, roots, k, scores);
}

void orc_brandes_f64(int64_t n, const int64_t* off, const int32_t* dst, const int32_t* roots, int k,
                     double* scores) {
  Brandes<double>(n, off, dst, roots, k, scores);
}

uint64_t orc_ordered_count(int64_t n, const int64_t* off, const int32_t* dst, double* seconds) {
  return orc_ordered_count_sample(n, off, dst, 1, 0, seconds);
}

// OrderedCount's outer loop restricted to u = first, first + stride, ...:
// the counts over first = 0 .. stride-1 sum to the full count, and one of
// them times a bounded sample of the whole run.
uint64_t orc_ordered_count_sample(int64_t n, const int64_t* off, const int32_t* dst, int64_t stride,
                                  int64_t first, double* seconds) {
  const auto t0 = Clock::now();
  uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64)
  for (int64_t u = first; u < n; u += stride) {
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = dst[e];
      if (v > u) break;
      int64_t it = off[v];
      for (int64_t f = off[u]; f < off[u + 1]; f++) {
        const int32_t w = dst[f];
        if (w > v) break;
        while (it < off[v + 1] && dst[it] < w) it++;
        if (it < off[v + 1] && dst[it] == w) total++;
      }
    }
  }
  if (seconds) *seconds = Since(t0);
  return total;
}

uint64_t orc_tc_intersect(int64_t n, const int64_t* off, const int32_t* dst) {
  uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64)
  for (int64_t u = 0; u < n; u++)
    for (int64_t e = off[u]; e < off[u + 1]; e++) {
      const int32_t v = dst[e];
      int64_t a = off[u], b = off[v];
      while (a < off[u + 1] && b < off[v + 1]) {
        if (dst[a] < dst[b]) {
          a++;
        } else if (dst[b] < dst[a]) {
          b++;
        } else {
          total++;
          a++;
          b++;
        }
      }
    }
  return total / 6;
}

int orc_depth_from_parent(int64_t n, const int32_t* parent, int32_t src, int32_t* depth) {
  std::fill(depth, depth + n, -2);
  if (parent[src] != src) return -1;
  depth[src] = 0;
  std::vector<int32_t> chain;
  for (int64_t v = 0; v < n; v++) {
    if (depth[v] != -2) continue;
    if (parent[v] < 0) {
      depth[v] = -1;
      continue;
    }
    chain.clear();
    int64_t x = v;
    while (depth[x] == -2) {
      if (parent[x] < 0 || parent[x] >= n || static_cast<int64_t>(chain.size()) > n) return -1;
      chain.push_back(static_cast<int32_t>(x));
      depth[x] = -3;  // on the current chain
      x = parent[x];
    }
    if (depth[x] < 0) return -1;  // cycle or a chain into an unreached vertex
    int32_t d = depth[x];
    for (auto it = chain.rbegin(); it != chain.rend(); ++it) depth[*it] = ++d;
  }
  return 0;
}

}  // extern "C"
