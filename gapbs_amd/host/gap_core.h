// Core host-side types and helpers shared by the graph builder, the drivers
// and the C-ABI shims.
//
// Reference counterparts:
//   NodeID / WeightT typedefs ............ src/benchmark.h:31-36
//   kRandSeed ............................. src/util.h:22
//   pvector (uninitialised array) ......... src/pvector.h:22-170
//   Timer ................................. src/timer.h:18-47
//   PrintLabel / PrintTime / PrintStep .... src/util.h:25-46
//   ParallelPrefixSum ..................... src/builder.h:105-137
//
// The print formats are kept byte-identical to the reference's because the
// reference's smoke tests grep those lines (test/test.mk:44-49, 76-79).
#pragma once

#include <omp.h>

#include <chrono>
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

namespace gap {

using NodeID = int32_t;
using WeightT = int32_t;
using Offset = int64_t;

// Seed of every random stream in the reference suite (generator, source
// picker).  Generated graphs may override it with a per-graph seed.
constexpr int64_t kRandSeed = 27491095;

// Host errors carry the exit status the reference uses for the same
// condition; drivers turn them into std::exit(code), the C-ABI into a
// negative return value.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Move-only, uninitialised, 64-byte aligned heap array.  Large graphs are
// tens of GB, so nothing is value-initialised unless asked for.
template <class T>
class Buf {
 public:
  Buf() = default;
  explicit Buf(size_t n) { reset(n); }
  Buf(size_t n, T v) {
    reset(n);
    fill(v);
  }
  ~Buf() { std::free(p_); }
  Buf(Buf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      std::free(p_);
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;

  void reset(size_t n) {
    std::free(p_);
    p_ = nullptr;
    n_ = n;
    if (n) {
      size_t bytes = (n * sizeof(T) + 63) & ~size_t(63);
      p_ = static_cast<T*>(std::aligned_alloc(64, bytes));
      if (!p_) throw Error(-40, "host allocation failed");
    }
  }
  void fill(T v) {
    const int64_t n = static_cast<int64_t>(n_);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p_[i] = v;
  }
  // Shrinks the logical size without reallocating.
  void truncate(size_t n) {
    if (n < n_) n_ = n;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T& operator[](size_t i) { return p_[i]; }
  const T& operator[](size_t i) const { return p_[i]; }
  T* begin() { return p_; }
  T* end() { return p_ + n_; }
  const T* begin() const { return p_; }
  const T* end() const { return p_ + n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

class Timer {
 public:
  void Start() { t0_ = t1_ = Clock::now(); }
  void Stop() { t1_ = Clock::now(); }
  double Seconds() const { return std::chrono::duration<double>(t1_ - t0_).count(); }
  double Millisecs() const { return Seconds() * 1e3; }

 private:
  using Clock = std::chrono::steady_clock;
  Clock::time_point t0_, t1_;
};

inline void PrintLabel(const std::string& label, const std::string& val) {
  std::printf("%-21s%7s\n", (label + ":").c_str(), val.c_str());
}

inline void PrintTime(const std::string& label, double seconds) {
  std::printf("%-21s%3.5lf\n", (label + ":").c_str(), seconds);
}

inline void PrintStep(const std::string& label, int64_t count) {
  std::printf("%-14s%14" PRId64 "\n", (label + ":").c_str(), count);
}

inline void PrintStep(const std::string& label, double seconds, int64_t count = -1) {
  if (count != -1)
    std::printf("%5s%11" PRId64 "  %10.5lf\n", label.c_str(), count, seconds);
  else
    std::printf("%5s%23.5lf\n", label.c_str(), seconds);
}

inline void PrintStep(int step, double seconds, int64_t count = -1) {
  PrintStep(std::to_string(step), seconds, count);
}

// Exclusive prefix sum of n counts into n+1 offsets, chunked per thread.
template <class CountT>
Buf<Offset> PrefixSum(const CountT* counts, int64_t n) {
  Buf<Offset> out(static_cast<size_t>(n) + 1);
  const int nthreads = omp_get_max_threads();
  Buf<Offset> partial(static_cast<size_t>(nthreads) + 1, 0);
  int used = 1;
#pragma omp parallel num_threads(nthreads)
  {
    const int t = omp_get_thread_num();
    const int nt = omp_get_num_threads();
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    Offset s = 0;
    for (int64_t i = lo; i < hi; i++) s += counts[i];
    partial[t + 1] = s;
#pragma omp barrier
#pragma omp single
    {
      used = nt;
      for (int i = 0; i < nt; i++) partial[i + 1] += partial[i];
    }
    Offset run = partial[t];
    for (int64_t i = lo; i < hi; i++) {
      out[i] = run;
      run += counts[i];
    }
  }
  out[n] = partial[used];
  return out;
}

}  // namespace gap
