// Feeds per-root times / traversed edges (stdin: n, then n lines of
// "bfs_time validate_time nedge") to the result-block printer.
#include <cstdio>
#include <vector>

#include "../../graph500-bfs_amd/host/report.hpp"

int main() {
  int n = 0;
  if (std::scanf("%d", &n) != 1) return 1;
  std::vector<double> t(n), v(n), e(n);
  for (int i = 0; i < n; ++i)
    if (std::scanf("%lf %lf %lf", &t[i], &v[i], &e[i]) != 3) return 1;
  g500host::print_bfs_result(t, v, e, true);
  return 0;
}
