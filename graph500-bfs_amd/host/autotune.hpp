// autotune.hpp -- root selection with the redraw rule and the adaptive
// alpha/beta tuner of the reference driver (mpi/main.cc:37-391), restated
// over an abstract runner so the same logic drives the GPU engine and the
// CPU unit test.
//
// Runner must provide
//   double run(int64_t root, double alpha, double beta, int64_t* at_data);
//          one BFS, returns seconds, fills at_data[9]
//   int find_roots(int64_t* roots, int n, int r1, int r2);
#pragma once

#include <algorithm>
#include <cinttypes>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <vector>

namespace g500host {

enum {
  kAtGlobalNqEdges = 0,
  kAtNumGlobalEdges = 1,
  kAtT2BLevel = 2,
  kAtGlobalNqSize = 3,
  kAtNumGlobalVerts = 4,
  kAtB2TLevel = 5,
  kAtPreGlobalNqEdges = 6,
  kAtPreGlobalNqSize = 7,
  kAtLevel = 8,
  kAtNum = 9
};
constexpr int64_t kAtNotDefined = -1;
constexpr int kRemoveRootsScaleThreshold = 6;  // mpi/parameters.h:277
constexpr int kRemoveRootsLevelThreshold = 3;  // mpi/parameters.h:278

// Traversed edges per root for edgefactor 16 and seed (2,3): the table the
// reference uses as the TEPS numerator while tuning (mpi/utils.hpp:3688-3733).
inline int64_t pf_nedge(int scale) {
  static const int64_t t[] = {
      -1, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16383, 32767, 65535,
      131070, 262144, 524285, 1048570, 2097137, 4194250, 8388513, 16776976,
      33553998, 67108130, 134216177, 268432547, 536865258, 1073731075,
      2147462776, 4294927670, 8589858508, 17179724952, 34359466407,
      68718955183, 137437972330, 274876029861, 549752273512, 1099505021204,
      2199010857419, 4398023184983, 8796049485746, 17592103986723,
      35184218756936, 70368458081736, 140736950689206};
  if (scale < 0 || scale >= (int)(sizeof(t) / sizeof(t[0]))) return -1;
  return t[scale];
}

// mpi/main.cc:37-43
inline double calc_teps(int root_start, int n, const double* perf) {
  double tmp = 0;
  for (int i = root_start; i < n; ++i) tmp += 1.0 / perf[i];
  return (n - root_start) / tmp;
}

template <class Runner>
class AutoTuner {
 public:
  using AtRow = int64_t[kAtNum];

  AutoTuner(Runner& runner, int scale, int edgefactor, bool verbose = true)
      : r_(runner), scale_(scale), ef_(edgefactor), verbose_(verbose) {}

  // mpi/main.cc:313-331
  void measure_performance(int root_start, int n, const int64_t* roots,
                           AtRow* at, double alpha, double beta, double* perf) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < kAtNum; ++j) at[i][j] = kAtNotDefined;
    for (int i = root_start; i < n; ++i) {
      const double t = r_.run(roots[i], alpha, beta, at[i]);
      perf[i] = (double)pf_nedge(scale_) / t;
    }
  }

  // mpi/main.cc:334-359: redraw every root while any root reaches <= 3
  // levels (SCALE > 6 only).
  void find_roots_in_large_graph(int root_start, int n, int64_t* roots, AtRow* at,
                                 double alpha, double beta, double* perf,
                                 int seed = 0) {
    int r = seed;
    r_.find_roots(roots, n, r++, 0);
    measure_performance(root_start, n, roots, at, alpha, beta, perf);
    if (scale_ <= kRemoveRootsScaleThreshold) return;
    for (;;) {
      bool ok = true;
      for (int i = root_start; i < n; ++i)
        if (at[i][kAtLevel] <= kRemoveRootsLevelThreshold) ok = false;
      if (ok) break;
      log("Pick roots again");
      r_.find_roots(roots, n, r++, 0);
      measure_performance(root_start, n, roots, at, alpha, beta, perf);
    }
  }

  // mpi/main.cc:45-311: one tuning step for alpha (is_alpha) or beta.
  // Returns true while the parameter keeps improving.
  bool auto_tuning_each(bool is_alpha, int root_start, int n, const int64_t* roots,
                        AtRow* at, double* alpha, double* beta, double* perf) {
    struct Item {
      int64_t val;
      int idx;
    };
    std::vector<Item> s, l;
    double pre_param = 0, pre_teps = calc_teps(root_start, n, perf);
    logf("========= START AUTO TUNING FOR %s =========", is_alpha ? "ALPHA" : "BETA");
    logf("Alpha = %f, Beta = %f, Harmonic Mean = %.0f MTEPS", *alpha, *beta,
         pre_teps / 1000000);
    for (int i = root_start; i < n; ++i) {
      int64_t tmp = is_alpha ? at[i][kAtGlobalNqEdges] : at[i][kAtPreGlobalNqSize];
      if (tmp != kAtNotDefined) s.push_back({tmp, i});
      tmp = is_alpha ? at[i][kAtPreGlobalNqEdges] : at[i][kAtGlobalNqSize];
      if (tmp != kAtNotDefined) l.push_back({tmp, i});
    }
    // std::sort by value only, as the reference sorts its auto_tuning_t
    // arrays (parameters.h:280-285): over more than 16 roots introsort
    // reorders equal values, and the same algorithm on the same sequence
    // reorders them the same way, so ties are nudged in the reference's order
    auto by_val = [](const Item& a, const Item& b) { return a.val < b.val; };
    std::sort(s.begin(), s.end(), by_val);
    std::sort(l.begin(), l.end(), by_val);

    bool smaller = false;
    if (s.empty()) return false;
    if (l.empty()) {
      smaller = true;
    } else if (perf[s[0].idx] < perf[l.back().idx]) {
      smaller = true;
    }
    logf("%s is %s", is_alpha ? "Alpha" : "Beta", smaller ? "smaller" : "larger");

    auto step = [&](const Item& it, bool decrease) -> int {
      // returns 0 = continue, 1 = stop (keep previous), 2 = stop (cannot move)
      pre_param = is_alpha ? *alpha : *beta;
      const int64_t* a = at[it.idx];
      const int pre_level = (int)(is_alpha ? a[kAtT2BLevel] : a[kAtB2TLevel]);
      const bool blocked =
          decrease ? ((is_alpha && pre_level + 1 == a[kAtLevel]) ||
                      (!is_alpha && pre_level == 1))
                   : ((is_alpha && pre_level == 1) ||
                      (!is_alpha && pre_level + 1 == a[kAtLevel]));
      if (blocked || pre_level == kAtNotDefined) {
        logf("[%02d] %s cannot be %s; %s is determined to be %f", it.idx,
             is_alpha ? "T2B" : "B2T",
             (decrease == is_alpha) ? "larger" : "smaller",
             is_alpha ? "Alpha" : "Beta", pre_param);
        return 2;
      }
      const double f = decrease ? 0.99 : 1.01;
      if (is_alpha) {
        const int64_t m = a[kAtNumGlobalEdges];
        *alpha = (double)m / it.val * f;
        if (decrease)
          while (it.val > (int64_t)(m / *alpha)) *alpha *= 0.99;
        else
          while (it.val <= (int64_t)(m / *alpha)) *alpha *= 1.01;
      } else {
        const int64_t nv = a[kAtNumGlobalVerts];
        *beta = (double)nv / (it.val * ef_ * 2.0) * f;
        if (decrease)
          while (it.val >= (int64_t)(nv / (*beta * ef_ * 2.0))) *beta *= 0.99;
        else
          while (it.val < (int64_t)(nv / (*beta * ef_ * 2.0))) *beta *= 1.01;
      }
      logf("[%02d] %s: %f -> %f", it.idx, is_alpha ? "Alpha" : "Beta", pre_param,
           is_alpha ? *alpha : *beta);
      for (int j = 0; j < kAtNum; ++j) at[it.idx][j] = kAtNotDefined;
      const double t = r_.run(roots[it.idx], *alpha, *beta, at[it.idx]);
      const double new_perf = (double)pf_nedge(scale_) / t;
      const int cur_level =
          (int)(is_alpha ? at[it.idx][kAtT2BLevel] : at[it.idx][kAtB2TLevel]);
      if (pre_level == cur_level)
        throw std::runtime_error("auto tuning: switch level did not move");
      const double pre_perf = perf[it.idx];
      perf[it.idx] = new_perf;
      const double new_teps = calc_teps(root_start, n, perf);
      logf("[%02d] %s: %d -> %d, Perf.: %.0f -> %.0f MTEPS (Harmonic Mean: %.0f -> %.0f MTEPS)",
           it.idx, is_alpha ? "T2B" : "B2T", pre_level, cur_level, pre_perf / 1e6,
           new_perf / 1e6, pre_teps / 1e6, new_teps / 1e6);
      if (pre_perf > new_perf) {
        perf[it.idx] = pre_perf;
        if (is_alpha)
          *alpha = pre_param;
        else
          *beta = pre_param;
        logf("%s is determined to be %f", is_alpha ? "Alpha" : "Beta", pre_param);
        return 1;
      }
      pre_teps = new_teps;
      return 0;
    };

    if (smaller) {
      // decreasing alpha raises T2B; decreasing beta lowers B2T
      for (size_t i = 0; i < s.size(); ++i)
        if (step(s[i], true)) return false;
    } else {
      for (size_t i = l.size(); i-- > 0;)
        if (step(l[i], false)) return false;
    }
    return true;
  }

  // mpi/main.cc:361-391
  double auto_tuning(int root_start, int n, int seed, int64_t* roots, double* alpha,
                     double* beta) {
    std::vector<AtRow> at((size_t)n);
    std::vector<double> perf((size_t)n);
    find_roots_in_large_graph(root_start, n, roots, at.data(), *alpha, *beta,
                              perf.data(), seed);
    while (auto_tuning_each(true, root_start, n, roots, at.data(), alpha, beta, perf.data()))
      measure_performance(root_start, n, roots, at.data(), *alpha, *beta, perf.data());
    while (auto_tuning_each(false, root_start, n, roots, at.data(), alpha, beta, perf.data()))
      measure_performance(root_start, n, roots, at.data(), *alpha, *beta, perf.data());
    const double teps = calc_teps(root_start, n, perf.data());
    logf("Estimated Performance = %.0f MTEPS with Alpha = %f and Beta = %f",
         teps / 1000000, *alpha, *beta);
    return teps;
  }

 private:
  void log(const char* s) {
    if (verbose_) std::printf("%s\n", s);
  }
  __attribute__((format(printf, 2, 3))) void logf(const char* fmt, ...) {
    if (!verbose_) return;
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
    std::printf("\n");
  }

  Runner& r_;
  int scale_, ef_;
  bool verbose_;
};

}  // namespace g500host
