// report.hpp -- the Graph500 result block (mpi/logfile.h:27-133).
//
// Order statistics use the same quartile rule as the reference:
// q1 = (x[(n-1)/4] + x[n/4]) / 2, etc.; harmonic mean TEPS = 1/mean(t/TE),
// and its standard error from Norris (1940) as the specification asks.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

namespace g500host {

struct Stats {
  double min, q1, median, q3, max, mean, std;
};

inline Stats statistics(std::vector<double> x) {
  Stats r{};
  const int n = (int)x.size();
  if (n == 0) return r;
  double s = 0;
  for (double v : x) s += v;
  r.mean = s / n;
  double v2 = 0;
  for (double v : x) v2 += (v - r.mean) * (v - r.mean);
  r.std = n > 1 ? std::sqrt(v2 / (n - 1)) : 0.0;
  std::sort(x.begin(), x.end());
  r.min = x[0];
  r.q1 = (x[(n - 1) / 4] + x[n / 4]) * .5;
  r.median = (x[(n - 1) / 2] + x[n / 2]) * .5;
  r.q3 = (x[n - 1 - (n - 1) / 4] + x[n - 1 - n / 4]) * .5;
  r.max = x[n - 1];
  return r;
}

inline void print_bfs_result(const std::vector<double>& bfs_times,
                             const std::vector<double>& validate_times,
                             const std::vector<double>& edge_counts, bool result_ok,
                             FILE* out = stdout) {
  if (!result_ok) {
    std::fprintf(out, "No results printed for invalid run.\n");
    return;
  }
  const int n = (int)bfs_times.size();
  Stats t = statistics(bfs_times);
  std::fprintf(out, "min_time:                       %.12g\n", t.min);
  std::fprintf(out, "firstquartile_time:             %.12g\n", t.q1);
  std::fprintf(out, "median_time:                    %.12g\n", t.median);
  std::fprintf(out, "thirdquartile_time:             %.12g\n", t.q3);
  std::fprintf(out, "max_time:                       %.12g\n", t.max);
  std::fprintf(out, "mean_time:                      %.12g\n", t.mean);
  std::fprintf(out, "stddev_time:                    %.12g\n", t.std);
  Stats e = statistics(edge_counts);
  std::fprintf(out, "min_nedge:                      %.11g\n", e.min);
  std::fprintf(out, "firstquartile_nedge:            %.11g\n", e.q1);
  std::fprintf(out, "median_nedge:                   %.11g\n", e.median);
  std::fprintf(out, "thirdquartile_nedge:            %.11g\n", e.q3);
  std::fprintf(out, "max_nedge:                      %.11g\n", e.max);
  std::fprintf(out, "mean_nedge:                     %.11g\n", e.mean);
  std::fprintf(out, "stddev_nedge:                   %.11g\n", e.std);
  std::vector<double> spe(n);
  for (int i = 0; i < n; ++i) spe[i] = bfs_times[i] / edge_counts[i];
  Stats p = statistics(spe);
  std::fprintf(out, "min_TEPS:                       %.12g\n", 1. / p.max);
  std::fprintf(out, "firstquartile_TEPS:             %.12g\n", 1. / p.q3);
  std::fprintf(out, "median_TEPS:                    %.12g\n", 1. / p.median);
  std::fprintf(out, "thirdquartile_TEPS:             %.12g\n", 1. / p.q1);
  std::fprintf(out, "max_TEPS:                       %.12g\n", 1. / p.min);
  std::fprintf(out, "harmonic_mean_TEPS:             %.12g\n", 1. / p.mean);
  std::fprintf(out, "harmonic_stddev_TEPS:           %.12g\n",
               n > 1 ? p.std / (p.mean * p.mean * std::sqrt((double)(n - 1))) : 0.0);
  Stats v = statistics(validate_times);
  std::fprintf(out, "min_validate:                   %.12g\n", v.min);
  std::fprintf(out, "firstquartile_validate:         %.12g\n", v.q1);
  std::fprintf(out, "median_validate:                %.12g\n", v.median);
  std::fprintf(out, "thirdquartile_validate:         %.12g\n", v.q3);
  std::fprintf(out, "max_validate:                   %.12g\n", v.max);
  std::fprintf(out, "mean_validate:                  %.12g\n", v.mean);
  std::fprintf(out, "stddev_validate:                %.12g\n", v.std);
}

}  // namespace g500host
