// CPU test of the restated alpha/beta tuner (graph500-bfs_amd/host/autotune.hpp,
// mpi/main.cc:45-391).  A scripted runner replays the reference's direction
// rule with the oracle (orc_do_schedule) on fixed per-level frontier sizes
// and charges a fixed cost per level, so every tuner decision is traced by
// hand below: which root is nudged, to which alpha/beta, whether the change
// is kept, and when tuning stops.
#include <cmath>
#include <cstdio>
#include <tuple>
#include <vector>

#include "../../graph500-bfs_amd/host/autotune.hpp"
#include "../../oracle/g500_oracle.h"

using g500host::AutoTuner;

struct Call {
  int64_t root;
  double alpha, beta;
};

// Cost of one run: a top-down level l costs 1 us per edge of the previous
// level's frontier (1 us for level 1), a bottom-up level 200 us + 0.1 us per
// vertex of its frontier.
struct FakeRunner {
  std::vector<std::vector<int64_t>> ns, ne;  // per root: nq_size, nq_edges by level (index 0 unused)
  int64_t m = 0, nc = 0;
  int ef = 16;
  std::vector<Call> calls;
  double run(int64_t root, double alpha, double beta, int64_t* at) {
    int8_t dirs[64];
    const auto& s = ns[(size_t)root];
    const auto& e = ne[(size_t)root];
    const int k = orc_do_schedule(m, nc, ef, alpha, beta, 0, s.data(), e.data(),
                                  (int)s.size(), at, dirs, 64);
    calls.push_back({root, alpha, beta});
    double t = 0;
    for (int l = 0; l < k; ++l) {
      const int lev = l + 1;
      const double prev_edges = lev >= 2 && lev - 1 < (int)e.size() ? (double)e[lev - 1] : 1.0;
      const double cur = lev < (int)s.size() ? (double)s[lev] : 0.0;
      t += dirs[l] ? 1e-6 * prev_edges : 1e-6 * (200.0 + 0.1 * cur);
    }
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

static bool same_calls(const std::vector<Call>& got, const std::vector<Call>& want) {
  if (got.size() != want.size()) return false;
  for (size_t i = 0; i < got.size(); ++i)
    if (got[i].root != want[i].root || got[i].alpha != want[i].alpha ||
        got[i].beta != want[i].beta)
      return false;
  return true;
}

int main() {
  // --- A: alpha, smaller branch kept, then larger branch reverted.
  // m = 1000, alpha 10 -> threshold 100.  Root 0: level 1 TD (40 edges),
  // level 2 TD with 150 > 100 edges -> T2B = 2, then BU to level 5.
  // Root 1 never passes 100 edges (90, 100), stays top-down.
  // Times: root 0 = 1 + 40 + 220 + 210 + 200 = 671 us, root 1 = 271 us, so
  // perf[s[0] = root 0] < perf[l.back() = root 1]: alpha is "smaller".
  // Step: alpha = 1000/150*0.99 = 6.6 (threshold 151): root 0 switches at
  // level 3 instead, 1 + 40 + 150 + 210 + 200 = 561 us < 671: kept.
  // Next pass (after measuring both at 6.6): s = {root 0: 999}, l =
  // {root 1: 100, root 0: 150}; perf[root 0] == perf[root 0] is not
  // smaller -> "larger": alpha = 1000/150*1.01 (threshold 148) puts root 0's
  // switch back at level 2 (671 us > 561): reverted to 6.6, alpha tuning ends.
  // Beta: root 0 never switched back (B2T undefined) and is the only entry of
  // s -> blocked at once; beta stays 14.
  {
    FakeRunner r;
    r.m = 1000;
    r.nc = 448;
    r.ns = {{0, 5, 50, 200, 100, 0}, {0, 4, 30, 60, 70, 0}};
    r.ne = {{0, 40, 150, 999, 999, 0}, {0, 20, 60, 90, 100, 0}};
    AutoTuner<FakeRunner> tuner(r, 16, 16, false);
    int64_t roots[2];
    double alpha = 10.0, beta = 14.0;
    tuner.auto_tuning(0, 2, 0, roots, &alpha, &beta);
    const double a1 = 1000.0 / 150 * 0.99, a2 = 1000.0 / 150 * 1.01;
    CHECK(alpha == a1);
    CHECK(beta == 14.0);
    CHECK(same_calls(r.calls, {{0, 10.0, 14.0}, {1, 10.0, 14.0},   // roots
                               {0, a1, 14.0},                       // smaller, kept
                               {0, a1, 14.0}, {1, a1, 14.0},        // measure
                               {0, a2, 14.0}}));                    // larger, reverted
  }
  // --- B: beta, smaller branch reverted.  n_c = 89600: beta 14 -> threshold
  // 89600 / (14*16*2) = 200.  Root 0 switches back at level 4 (150 < 200),
  // root 1 at level 5 (frontier 1000 at level 4, then 100).  s = {root 1:
  // 1000} (pre size), l = {root 1: 100, root 0: 150}.  Times: root 0 =
  // 1 + 40 + 400 + 215 + 300 + 60 = 1016 us, root 1 = 1 + 30 + 500 + 300 +
  // 210 + 50 = 1091 us -> perf[root 1] < perf[root 0]: "smaller".  beta =
  // 89600 / (1000*16*2) * 0.99 = 2.772 (threshold 1010) moves root 1's switch
  // to level 4 (then a 2000-edge top-down level: 2881 us): reverted to 14.
  {
    FakeRunner r;
    r.m = 1000;
    r.nc = 89600;
    r.ns = {{0, 5, 50, 2000, 150, 40, 0}, {0, 3, 30, 3000, 1000, 100, 0}};
    r.ne = {{0, 40, 900, 5000, 300, 60, 0}, {0, 30, 700, 6000, 2000, 50, 0}};
    AutoTuner<FakeRunner> tuner(r, 16, 16, false);
    int64_t roots[2];
    int64_t at[2][9];
    double perf[2];
    double alpha = 10.0, beta = 14.0;
    tuner.find_roots_in_large_graph(0, 2, roots, at, alpha, beta, perf);
    CHECK(at[0][5] == 4 && at[0][3] == 150 && at[0][7] == -1 && at[0][8] == 6);
    CHECK(at[1][5] == 5 && at[1][3] == 100 && at[1][7] == 1000);
    const bool more = tuner.auto_tuning_each(false, 0, 2, roots, at, &alpha, &beta, perf);
    const double b1 = 89600.0 / (1000 * 16 * 2.0) * 0.99;
    CHECK(!more);
    CHECK(beta == 14.0);
    CHECK(same_calls(r.calls, {{0, 10.0, 14.0}, {1, 10.0, 14.0}, {1, 10.0, b1}}));
    CHECK(at[1][5] == 4);  // the reverted run's data stays, as in the reference
  }
  // --- C: beta, larger branch kept, then blocked.  As B, but root 0's level
  // 5 would scan 2000 edges top-down: root 0 = 2716 us, root 1 = 1091 us ->
  // "larger".  l from the end: root 0 (150): beta = 89600/(150*32)*1.01
  // (threshold 148) keeps root 0 bottom-up one more level (B2T 4 -> 5):
  // 1 + 40 + 400 + 215 + 204 + 60 = 920 us, kept.  Next l entry, root 1:
  // B2T 5 + 1 == its 6 levels -> "B2T cannot be larger": stop, beta kept.
  {
    FakeRunner r;
    r.m = 1000;
    r.nc = 89600;
    r.ns = {{0, 5, 50, 2000, 150, 40, 0}, {0, 3, 30, 3000, 1000, 100, 0}};
    r.ne = {{0, 40, 900, 5000, 2000, 60, 0}, {0, 30, 700, 6000, 2000, 50, 0}};
    AutoTuner<FakeRunner> tuner(r, 16, 16, false);
    int64_t roots[2];
    int64_t at[2][9];
    double perf[2];
    double alpha = 10.0, beta = 14.0;
    tuner.find_roots_in_large_graph(0, 2, roots, at, alpha, beta, perf);
    const bool more = tuner.auto_tuning_each(false, 0, 2, roots, at, &alpha, &beta, perf);
    const double b1 = 89600.0 / (150 * 16 * 2.0) * 1.01;
    CHECK(!more);
    CHECK(beta == b1);
    CHECK(same_calls(r.calls, {{0, 10.0, 14.0}, {1, 10.0, 14.0}, {0, 10.0, b1}}));
    CHECK(at[0][5] == 5);
  }
  // --- D: ties over more than 16 roots follow std::sort, as in the
  // reference.  20 roots, each T2B = 2 after a first level of v_i frontier
  // edges (v below), so l = the v_i in root order and every root costs
  // (401.4 + v_i) us.  The branch is "larger" (no root is slower than one
  // with v = 5), so the first root nudged is the last entry of l among the
  // seven roots with v = 5: std::sort (introsort) leaves root 2 there, where
  // a stable sort would leave root 19.  alpha = 1000/5*1.01 = 202 moves its
  // switch to level 1 (601.7 us): reverted.
  {
    FakeRunner r;
    r.m = 1000;
    r.nc = 448;
    const int64_t v[20] = {5, 3, 5, 1, 5, 3, 2, 5, 3, 1, 5, 2, 3, 5, 1, 2, 5, 3, 2, 5};
    for (int i = 0; i < 20; ++i) {
      r.ns.push_back({0, 2, 3, 4, 0});
      r.ne.push_back({0, v[i], 500, 0, 0});
    }
    AutoTuner<FakeRunner> tuner(r, 16, 16, false);
    int64_t roots[20];
    int64_t at[20][9];
    double perf[20];
    double alpha = 10.0, beta = 14.0;
    tuner.find_roots_in_large_graph(0, 20, roots, at, alpha, beta, perf);
    CHECK(at[7][2] == 2 && at[7][6] == 5 && at[7][0] == 500);
    const bool more = tuner.auto_tuning_each(true, 0, 20, roots, at, &alpha, &beta, perf);
    CHECK(!more);
    CHECK(alpha == 10.0);
    CHECK(r.calls.size() == 21);
    CHECK(r.calls.size() == 21 && r.calls[20].root == 2 && r.calls[20].alpha == 1000.0 / 5 * 1.01);
  }
  // --- redraw rule: roots reaching <= 3 levels are drawn again (SCALE > 6)
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
    r.ns = {{0, 5, 50, 200, 0}, {0, 4, 30, 60, 0}, {0, 3, 0}};
    r.ne = {{0, 40, 150, 999, 0}, {0, 20, 60, 90, 0}, {0, 10, 0}};
    AutoTuner<Shallow> tuner(r, 16, 16, false);
    int64_t roots[2];
    int64_t at[2][9];
    double perf[2];
    tuner.find_roots_in_large_graph(0, 2, roots, at, 10.0, 14.0, perf);
    CHECK(r.draws == 2);  // root 2 has 2 levels -> redraw once
    CHECK(roots[1] == 1);
    // at SCALE <= 6 the rule is off
    Shallow r6;
    r6.m = r.m;
    r6.nc = r.nc;
    r6.ns = r.ns;
    r6.ne = r.ne;
    AutoTuner<Shallow> t6(r6, 6, 16, false);
    t6.find_roots_in_large_graph(0, 2, roots, at, 10.0, 14.0, perf);
    CHECK(r6.draws == 1 && roots[1] == 2);
  }
  if (failures) return 1;
  std::printf("autotune ok\n");
  return 0;
}
