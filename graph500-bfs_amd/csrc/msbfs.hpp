// msbfs.hpp -- handle of the batched bit-parallel multi-source BFS.
//
// Up to 256 roots run together: every vertex carries K = ceil(roots/64)
// 64-bit lane masks (bit i of word k = root 64k+i).  Level l's frontier is
// kept as its own array F_l (vertex-major, K words per vertex, by new ID), so
// the per-root levels can be read back after the traversal without any
// per-(vertex, root) writes during it.  A one-bit-per-vertex bitmap of each
// frontier (fits in L2) lets sweeps skip the lane words of vertices that are
// in no lane's frontier.
#pragma once

#include <vector>

#include "internal.hpp"

struct g500_msbfs {
  g500_graph* g = nullptr;
  hipStream_t stream = nullptr;
  int max_roots = 0;
  int kmax = 0;  // words per vertex at max_roots

  // per-run state
  int nroots = 0;
  int K = 0;
  bool in_flight = false;
  bool have_result = false;

  g500::DevBuf<uint64_t> seen;              // n_c * kmax
  std::vector<g500::DevBuf<uint64_t>*> F;   // level frontiers, n_c * kmax each
  g500::DevBuf<uint64_t*> fptr;             // device table of F pointers
  g500::DevBuf<unsigned long long> cnt;     // lane discoveries per level [cap + 2]
  g500::DevBuf<unsigned long long> te;      // per root
  g500::DevBuf<uint32_t> cnt0;              // input edges by source, by new ID (n)
  g500::DevBuf<int64_t> roots;              // device copy of the batch's roots
  int64_t* h_roots = nullptr;               // pinned staging for the roots
  unsigned long long* h_cnt = nullptr;      // pinned copy of cnt after a run
  g500::DevBuf<uint8_t> out;                // levels [max_roots][n] by original ID
  g500::DevBuf<uint64_t> planes;            // with out: levels as 4 bit planes per vertex (K <= 2)
  bool level_arrays = true;                 // false: point lookups only (no `out`)
  g500::DevBuf<uint32_t> stream_tile_v0;    // streaming pull: first vertex per 1024 edges
  g500::DevBuf<int64_t> gather_pairs;       // point-lookup staging
  g500::DevBuf<uint8_t> gather_levels;

  g500::DevBuf<uint64_t> fb[2];             // frontier bitmaps (level & 1), n_c bits
  g500::DevBuf<unsigned long long> fe;      // frontier degree sum per level [cap + 2]
  g500::DevBuf<uint64_t> act;               // lanes present in each level's frontier [(cap + 2) * 4]

  g500_msbfs_tuning tuning{};
  int64_t bin_heavy = 0, bin_medium = 0;    // new-ID bin ends for the tuning
  unsigned long long td_threshold = 0;      // top-down while frontier edges < this
  g500::DevBuf<uint64_t> item_start;        // heavy work items: prefix by vertex (h+1)
  g500::DevBuf<uint32_t> item_vertex;       // owner of each heavy work item
  std::vector<uint64_t> h_item_start;       // host copy of item_start
  int64_t n_items = 0;
  g500::DevBuf<uint64_t> acc;               // heavy pull accumulators (h * kmax)
  g500::DevBuf<unsigned long long> lane_part;  // split level: per-block lane degree sums
  g500::DevBuf<uint64_t> split_mask;        // split level: the lanes that push (4 words)
  g500::ScanWorkspace scan_ws;

  int run_levels = 0;   // level buffers the running batch's steps fill
  int depth_hint = 0;   // levels of the deepest batch so far, plus two (0: none yet)
  // one instantiated graph per K (the level loop is captured once per K)
  hipGraphExec_t exec[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int exec_levels[5] = {0, 0, 0, 0, 0};

  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  g500_msbfs_stats stats{};

  int levels_cap() const { return (int)F.size(); }
  ~g500_msbfs();
};

namespace g500 {
void msbfs_init(g500_msbfs* b, int max_roots, bool level_arrays);
// device bytes msbfs_init allocates (before the heavy work-item tables)
int64_t msbfs_footprint(const g500_graph* g, int max_roots, bool level_arrays);
void msbfs_set_tuning(g500_msbfs* b, const g500_msbfs_tuning& t);
void msbfs_launch(g500_msbfs* b, const int64_t* roots, int nroots);
bool msbfs_query(g500_msbfs* b);
void msbfs_wait(g500_msbfs* b);
void msbfs_gather(g500_msbfs* b, const int64_t* pairs, int64_t count, uint8_t* levels);
// the same lookups staged in a file-scope device array (no allocation)
void msbfs_gather_noalloc(g500_msbfs* b, const int64_t* pairs, int64_t count,
                          uint8_t* levels);
double msbfs_autotune(g500_msbfs* b, const int64_t* roots, int nroots, int rounds,
                      g500_msbfs_tuning* best);
}  // namespace g500
