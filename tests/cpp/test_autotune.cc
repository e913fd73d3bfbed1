// CPU test of the restated alpha/beta tuner (graph500-bfs_amd/host/autotune.hpp).
// A synthetic runner replays the reference's direction decisions with the
// oracle (orc_do_schedule) on fixed per-level frontier sizes and charges a
// simple cost per level, so every tuner decision can be checked by hand.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../graph500-bfs_amd/host/autotune.hpp"
#include "../../oracle/g500_oracle.h"

using g500host::AutoTuner;

struct FakeRunner {
  std::vector<std::vector<int64_t>> ns, ne;  // per root: nq_size, nq_edges by level
  int64_t m, nc;
  int ef;
  int runs = 0;
  std::vector<std::vector<int>> dirs_log;
  double run(int64_t root, double alpha, double beta, int64_t* at) {
    int8_t dirs[64];
    const auto& s = ns[(size_t)root];
    const auto& e = ne[(size_t)root];
    const int k = orc_do_schedule(m, nc, ef, alpha, beta, 0, s.data(), e.data(),
                                  (int)s.size(), at, dirs, 64);
    ++runs;
    // cost: a top-down level scans the previous frontier's edges, a
    // bottom-up level costs a fixed sweep plus its frontier
    double t = 0;
    std::vector<int> d;
    for (int l = 0; l < k; ++l) {
      d.push_back(dirs[l]);
      const int lev = l + 1;
      const double prev_edges = lev >= 2 && lev - 1 < (int)e.size() ? (double)e[lev - 1] : 1.0;
      const double cur = lev < (int)s.size() ? (double)s[lev] : 0.0;
      t += dirs[l] ? 1e-6 * prev_edges : 1e-6 * (200.0 + 0.1 * cur);
    }
    dirs_log.push_back(d);
    return t;
  }
  int find_roots(int64_t* roots, int n, int, int) {
    for (int i = 0; i < n; ++i) roots[i] = i;
    return n;
  }
};

static int failures = 0;
#define CHECK(c)                                                    \
  do {                                                              \
    if (!(c)) {                                                     \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);      \
      ++failures;                                                   \
    }                                                               \
  } while (0)

int main() {
  // --- hand-traced single step: alpha moves just past the T2B frontier
  {
    FakeRunner r;
    r.m = 1000;
    r.nc = 448;
    r.ef = 16;
    // root 0 switches at level 2 (nq_edges 150 > 1000/10); root 1 never
    // grows past the threshold (its frontier edges stay <= 100)
    r.ns = {{0, 5, 50, 200, 100, 0}, {0, 4, 30, 60, 70, 0}};
    r.ne = {{0, 40, 150, 999, 999, 0}, {0, 20, 60, 90, 100, 0}};
    AutoTuner<FakeRunner> tuner(r, 16, 16, false);
    int64_t roots[2];
    int64_t at[2][9];
    double perf[2];
    double alpha = 10.0, beta = 14.0;
    tuner.find_roots_in_large_graph(0, 2, roots, at, alpha, beta, perf);
    CHECK(at[0][2] == 2 && at[0][0] == 150);  // T2B at level 2
    CHECK(at[1][2] == -1 && at[1][6] == 100);  // never switched
    const bool more = tuner.auto_tuning_each(true, 0, 2, roots, at, &alpha, &beta, perf);
    // s = {root0 (150)}, l = {root1 (100)}; which branch depends on perf
    if (perf[0] < perf[1]) {
      // decreasing alpha: 1000/150*0.99 = 6.6 -> threshold 151 >= 150
      CHECK(std::fabs(alpha - 6.6) < 1e-9 || alpha == 10.0);
    }
    (void)more;
  }
  // --- full tuning terminates and never lowers the harmonic mean
  {
    FakeRunner r;
    r.m = 100000;
    r.nc = 20000;
    r.ef = 16;
    for (int i = 0; i < 8; ++i) {
      const int64_t a = 1 + i, b = 20 * (i + 2), c = 2000 + 300 * i, d = 9000 - 500 * i,
                    e = 600 + 50 * i, f = 3 + i;
      r.ns.push_back({0, a, b, c, d, e, f, 0});
      r.ne.push_back({0, 60 * a, 9000 + 1000 * i, 50000, 40000, 3000, 20, 0});
    }
    AutoTuner<FakeRunner> tuner(r, 20, 16, false);
    int64_t roots[8];
    double alpha = 10.0, beta = 14.0;
    std::vector<int64_t[9]> at(8);
    std::vector<double> perf(8);
    tuner.find_roots_in_large_graph(0, 8, roots, at.data(), alpha, beta, perf.data());
    const double before = g500host::calc_teps(0, 8, perf.data());
    double ta = 10.0, tb = 14.0;
    const double after = tuner.auto_tuning(0, 8, 0, roots, &ta, &tb);
    CHECK(after >= before * (1 - 1e-12));
    CHECK(ta > 0 && tb > 0);
    CHECK(r.runs < 2000);
    std::printf("tuning: %.3f -> %.3f (alpha %.4f beta %.4f, %d runs)\n", before, after,
                ta, tb, r.runs);
  }
  // --- redraw rule: roots reaching <= 3 levels are drawn again
  {
    struct Shallow : FakeRunner {
      int draws = 0;
      int find_roots(int64_t* roots, int n, int r1, int) {
        ++draws;
        for (int i = 0; i < n; ++i) roots[i] = (r1 == 0 && i == 1) ? 2 : i;
        return n;
      }
    } r;
    r.m = 1000;
    r.nc = 448;
    r.ef = 16;
    r.ns = {{0, 5, 50, 200, 0}, {0, 4, 30, 60, 0}, {0, 3, 0}};
    r.ne = {{0, 40, 150, 999, 0}, {0, 20, 60, 90, 0}, {0, 10, 0}};
    AutoTuner<Shallow> tuner(r, 16, 16, false);
    int64_t roots[2];
    int64_t at[2][9];
    double perf[2];
    tuner.find_roots_in_large_graph(0, 2, roots, at, 10.0, 14.0, perf);
    CHECK(r.draws == 2);  // root 2 has 2 levels -> redraw once
    CHECK(roots[1] == 1);
  }
  if (failures) return 1;
  std::printf("autotune ok\n");
  return 0;
}
