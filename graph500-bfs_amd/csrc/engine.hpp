// engine.hpp -- the BFS engine handle (one per concurrent BFS stream).
#pragma once

#include <vector>

#include "internal.hpp"

// Steps of the level loop.  A step is one level (top-down, a top-down level
// run as a bottom-up sweep plus bitmap -> queue, or bottom-up over the bitmap
// or the unvisited lists) or the bitmap -> queue conversion at the
// bottom-up -> top-down switch.  Its code packs the level number and the mode.
enum StepMode : uint32_t {
  kModeTD = 1,       // top-down, frontier prepared by the prefix-sum kernels
  kModeTDX = 2,      // top-down level run as a bottom-up sweep + bitmap -> queue
  kModeBUSweep = 3,  // bottom-up over the bitmap
  kModeBUList = 4,   // bottom-up over the unvisited lists
  kModeConv = 5,     // bitmap -> queue at the bottom-up -> top-down switch
  kModeDone = 6,
  kModeTDS = 7,      // top-down, small frontier staged by the expansion itself
};
__host__ __device__ inline bool mode_is_td(uint32_t m) {
  return m == kModeTD || m == kModeTDS || m == kModeTDX;
}
__host__ __device__ inline uint32_t step_code(int level, uint32_t mode) {
  return ((uint32_t)level << 4) | mode;
}
__host__ __device__ inline uint32_t step_mode(uint32_t code) { return code & 15u; }
__host__ __device__ inline int step_level(uint32_t code) { return (int)(code >> 4); }

// Device-side state of the level loop.  cnt[] are the counters that queue
// reservations take within a level (top-down next-queue slots in cnt[0],
// long-row slots in cnt[2]); the step's closing kernel adds them to the
// blocks' totals and clears them.  The closing kernel also applies the
// direction rule and writes the code of the step that must run next into
// `step`; every kernel of a step compares that code with its own and returns
// at once when they differ.  So the host may queue a predicted step behind
// the running one: a wrong prediction costs its empty launches, never a
// result.  The per-root init kernel sets every field.
struct LevelSync {
  unsigned long long cnt[4];
  uint32_t step;       // code of the step to run next
  uint32_t growing;    // the frontier has grown at every level so far
  long long qlen;      // top-down frontier queue length
  long long qedges;    // its degree sum (>= its CSR entries)
  long long prev_nq;   // frontier size of the previous level
  long long alpha_thr, beta_thr, exec_thr;  // int64(m/alpha), int64(n_c/(beta*ef*2)), m/exec_alpha
  uint32_t snap_tag;   // code of the conversion that wrote the gather's bitmap snapshot (0: none)
};

// Host-visible (pinned, coherent, mapped) step reports: a ring of slots, the
// step with sequence number s posting to slot s % kReportSlots, so a step
// queued ahead of the host never overwrites a report the host has not read.
// Each word carries its value above a 16-bit tag (the low bits of s), so a
// report is whole once every word carries the tag: no word needs to wait for
// another (no system-scope fence on the device).
struct HostReport {
  unsigned long long w[4];  // frontier size, frontier degree sum, next step code, 0 (<< 16 | tag)
};
constexpr int kReportSlots = 8;
constexpr int kReportTagBits = 16;

struct g500_bfs {
  g500_graph* g = nullptr;
  hipStream_t stream = nullptr;

  int64_t words = 0;                // bitmap words covering [0, n_c)
  g500::DevBuf<int64_t> pred;       // n, by original ID (packed entries)
  g500::DevBuf<int64_t> pred_new;   // n_c, the same entries by new ID
  g500::DevBuf<uint64_t> vis, vis_alt;  // visited bitmaps (new IDs)
  g500::DevBuf<uint32_t> q, nq;     // frontier queues (new IDs)
  g500::DevBuf<uint64_t> dsum;      // frontier degree prefix sums
  g500::DevBuf<uint32_t> tile_owner; // frontier index owning each top-down tile's first edge
  g500::DevBuf<uint64_t> bsum, boff;  // per prep block: degree total, its offset
  g500::DevBuf<unsigned int> prep_ticket;  // blocks done in the prep's first launch
  // unvisited vertices after a bottom-up level: 4096 entries per chunk of 64
  // bitmap words, unv_cnt[chunk] of them in use
  g500::DevBuf<uint32_t> unv, unv_cnt;
  unsigned bu_grid[3] = {1, 1, 1};  // resident blocks of the sweep / list / long-row kernels
  g500::DevBuf<LevelSync> lsync;   // level-loop state on the device
  HostReport* h_report = nullptr;  // kReportSlots slots of pinned host memory, written by the GPU
  HostReport* d_report = nullptr;  // their device address
  unsigned long long seq = 0;      // last report sequence number issued
  // guessed steps queued behind the running one (G500_SPECULATE=<n>; 0: the
  // host waits for every step's report before it launches the next)
  int spec_depth = 1;
  // test hook (G500_GUESS=wrong): every guess is turned into a wrong one of
  // another kind, so the guarded launches, the guessed gather and the
  // guessed fixup run against a loop state that does not name them
  bool guess_wrong = false;
  int64_t unreach = 0;  // vertices with edges the last run left unreached

  // validation scratch
  g500::DevBuf<unsigned long long> vcount;  // [G500_VERR_NUM + 2]
  g500::DevBuf<uint8_t> pred_valid;         // n

  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  // the pred gather of the vertices reached by the last bottom-up level runs
  // on gstream during the remaining top-down levels; vis_snap is the bitmap
  // it covers (copied by the conversion's count pass), the fixup adds the rest
  hipStream_t gstream = nullptr;
  g500::DevBuf<uint64_t> vis_snap;
  // ev_snap follows the last conversion's count pass (the gather waits on
  // it); the engine stream waits on ev_g1, behind the last gather queued,
  // before the fixup and the next run.  Events, not value waits: a value
  // wait polls from a kernel, which never ends where the kernels of a device
  // run one at a time (as under a counter profiler)
  hipEvent_t ev_snap = nullptr;
  g500::DevBuf<uint64_t> lpart;   // per-block totals of a level's kernels
  g500::DevBuf<uint32_t> qcnt;    // per-wave counts of the bitmap -> queue conversion
  g500::DevBuf<uint64_t> qedges;  // its per-block degree sums
  // without CoreBFS: (original ID, new ID) of the vertices with edges, in
  // ascending original ID (the gather's domain; every other pred entry stays -1)
  g500::DevBuf<uint2> core;
  int64_t n_core = 0;
  int64_t iso_root = -1;  // a root without edges whose entry the last run wrote
  uint32_t epoch = 0;     // the run's epoch in bits 32-47 of its pred_new entries
  bool pred_fill_needed = false;  // pred must be reset to -1 before the next run
  std::vector<hipEvent_t> level_events;  // 2 per launched step when timing is on
  bool timing = false;
  // per-kernel timing (always on): the gather's events on the stream it runs
  // on (ev_g1 also orders the engine stream after the gather), and a pair
  // around each k_bu_sweep launch of a run
  hipEvent_t ev_g0 = nullptr, ev_g1 = nullptr;
  bool gather_pending = false;  // a gather on gstream neither stream nor host has waited for
  std::vector<hipEvent_t> sweep_events;  // 2 per launched sweep
  // counting runs: per-block row-scan bytes of the bottom-up kernels, one
  // slot per block (sweep, long-row and list kernels in that order)
  bool counting = false;
  g500::DevBuf<uint64_t> wacc;
  double exec_alpha = 25.0;  // > 0: top-down levels with > m/exec_alpha edges run bottom-up

  bool have_result = false;
  int64_t last_root = -1;
  g500_run_stats stats{};
  std::vector<g500_level_trace> trace;

  ~g500_bfs();
};

namespace g500 {
// allocates the engine; with CoreBFS the first engine of a graph also runs
// the reachability BFS with these settings (mpi/corebfs_adaptor.hpp:228-249)
void bfs_init(g500_bfs* b, int edgefactor, double alpha, double beta);
void bfs_run(g500_bfs* b, int64_t root, int edgefactor, double alpha,
             double beta, int64_t* at_data);
void bfs_validate(g500_bfs* b, int64_t root, const int64_t* host_pred,
                  int64_t* te, int64_t* errs, int* ok);
void bfs_levels(g500_bfs* b, uint16_t* out);
}  // namespace g500
