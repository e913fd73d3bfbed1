// bfs.hip -- direction-optimizing BFS on one MI355X.
//
// Level loop and alpha/beta switch follow BfsBase::run_bfs_core
// (mpi/bfs.hpp:3238-3566, decision at 3413-3457).  The data layout is our
// own: vertices renumbered by descending degree ("new" IDs), a uint64 visited
// bitmap over new IDs, uint32 frontier queues, and packed (parent | depth <<
// 48) entries written by new ID during the levels and gathered into pred by
// original ID at the end.
//
// Kernels
//   k_td_expand   top-down: one prefix sum over frontier degrees and a
//                 tile -> owner table (k_td_prep_sums, k_td_prep_tiles; a
//                 frontier of up to 512 vertices is staged by the expansion
//                 itself), then a persistent grid
//                 walks 2048-edge tiles; the tile's slice of the prefix sum
//                 sits in LDS and each lane binary-searches its edges' owners
//                 there (every lane gets eight edges, whatever the degree
//                 skew).  Winners of the visited-bit atomicOr write pred and
//                 are appended to an LDS queue that is flushed with one global
//                 atomic per tile.
//   k_bu_sweep    first bottom-up level after top-down ones: a wave owns 64
//                 bitmap words; lane l takes bit l of each word, so the
//                 first-neighbour probes read `first` coalesced, eight
//                 vertices per lane in flight.  The still-unvisited vertices
//                 go to per-chunk lists.
//   k_bu_list     later bottom-up levels: the same probes over those lists,
//                 compacted in place as vertices are found.  New bits are
//                 written once per bitmap word either way.
//   k_queue_count, k_queue_write   bitmap -> queue at the bottom-up ->
//                 top-down switch and after top-down levels run as sweeps.
//   k_pred_gather_core / k_pred_gather   pred[o] by original ID, gathered
//                 from the entries the levels wrote by new ID (levels write
//                 densely by new ID; one pass in original-ID order then
//                 produces pred).  Without CoreBFS it walks only the
//                 vertices with edges and starts on a second stream at the
//                 bottom-up -> top-down switch, against a bitmap snapshot;
//                 k_pred_fixup adds the vertices the last levels reach.
//   k_level_decide, k_queue_decide   one block closing each step: the
//                 step's totals, the direction rule, the next step's code.
//
// The level loop is a sequence of steps (a level, or the bitmap -> queue
// conversion at the bottom-up -> top-down switch).  Every block of a step's
// kernels writes its totals to a per-block slot; the one-block closing
// kernel sums the slots, applies the reference's direction rule on the
// device, writes the code of the step that must run next into the loop
// state and posts the totals with that code to pinned host memory.  Every
// kernel of a step first compares its own code with that state and returns
// at once when they differ.  So the host queues the step it predicts behind
// the running one before that one reports (no idle GPU between levels when
// the guess is right); when the posted code says otherwise, the queued
// guess returns at once and the host queues the right step.  The host also
// applies the rule to the posted totals (auto_tuning_data, the trace) and
// checks the device's choice against it.  No copies, memsets or stream
// synchronisations per level.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "engine.hpp"

namespace g500 {

namespace {

constexpr int kT = 256;
constexpr int kWavesPerBlock = kT / kWave;
constexpr int kTdItems = 8;
constexpr int kTdTile = kT * kTdItems;
constexpr int kTdWin = 512;  // owners staged in LDS per tile (more: read from memory)
constexpr int kMaxGrid = 2048;
// bits 32-47 of a level entry: the run's epoch (see run_entry); kEpochNone
// marks entries of no run
constexpr uint64_t kEpochMask = 0xFFFFull << 32;
constexpr uint32_t kEpochNone = 0xFFFF;

// The top-down step for a frontier of qlen vertices with degree sum qedges:
// run as a sweep plus conversion past the execution threshold, else with or
// without the prefix-sum preparation.
__host__ __device__ inline uint32_t td_step_mode(long long qlen, long long qedges,
                                                 long long exec_thr) {
  return qedges > exec_thr ? kModeTDX : (qlen <= kTdWin ? kModeTDS : kModeTD);
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Sum over a block of W waves; valid in thread 0.
template <int W = kWavesPerBlock>
__device__ inline uint64_t block_sum(uint64_t v, uint64_t* lds) {
  v = wave_sum_u64(v);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  uint64_t s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < W; ++w) s += lds[w];
  return s;
}

__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t* wsum,
                                           uint64_t* total) {
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  uint64_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint64_t y = __shfl_up(x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) wsum[wave] = x;
  __syncthreads();
  uint64_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kWavesPerBlock; ++w) {
    off += (w < wave) ? wsum[w] : 0;
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return x - v + off;
}

// Step totals to the host's pinned report ring (slot seq % kReportSlots),
// each word tagged with the low bits of seq.  The host needs nothing else
// the step wrote, so the words go out as plain system-scope stores: no
// release fence (which would write the L2 back) on the device's path.
__device__ inline void post_report(HostReport* ring, const unsigned long long v[4],
                                   unsigned long long seq) {
  HostReport* hr = ring + (seq % kReportSlots);
  const unsigned long long tag = seq & ((1ull << kReportTagBits) - 1);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    __hip_atomic_store(&hr->w[i], (v[i] << kReportTagBits) | tag, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// Every kernel of a step starts with this test: the step runs only if the
// level loop's state names it as the next one.  The closing kernel of the
// step before wrote that code (kernels of one stream see the writes of the
// kernels before them), and nothing writes it while the step runs.
__device__ inline bool step_is_next(const LevelSync* ls, uint32_t code) {
  return __builtin_amdgcn_readfirstlane(ls->step) == code;
}

// The direction rule of BfsBase::run_bfs_core (mpi/bfs.hpp:3413-3457) applied
// on the device after a level with totals (nq_size, nq_edges): top-down ->
// bottom-up when the frontier grew and its degree sum passes int64(m/alpha);
// bottom-up -> top-down (for good) once it stops growing and holds fewer than
// int64(n_c/(beta*ef*2)) vertices.  Picks the next step -- a top-down level
// with more than m/exec_alpha frontier edges runs as a sweep plus conversion
// (kModeTDX); a bottom-up level walks the unvisited lists unless a top-down
// expansion ran last; the switch back to top-down starts with the
// conversion -- writes its code and returns it.  One thread.
__device__ inline uint32_t decide_next(const LevelSync& cur, LevelSync* ls, uint32_t code,
                                       long long nq_size, long long nq_edges) {
  const uint32_t mode = step_mode(code);
  const bool td = mode_is_td(mode);
  uint32_t next;
  if (nq_size == 0) {
    next = kModeDone;
  } else {
    bool next_td = td, growing = cur.growing != 0;
    if (growing && nq_size > cur.prev_nq) {
      if (td && nq_edges > cur.alpha_thr) next_td = false;
    } else if (!td && nq_size < cur.beta_thr) {
      next_td = true;
      growing = false;
    }
    ls->growing = growing ? 1u : 0u;
    ls->prev_nq = nq_size;
    if (td && next_td) {
      next = td_step_mode(nq_size, nq_edges, cur.exec_thr);
      ls->qlen = nq_size;
      ls->qedges = nq_edges;
    } else if (td) {
      // a top-down expansion leaves the unvisited lists behind, a sweep
      // (also one that ran a top-down level) refreshes them
      next = mode == kModeTDX ? kModeBUList : kModeBUSweep;
    } else {
      next = next_td ? kModeConv : kModeBUList;
    }
  }
  const uint32_t nc = step_code(step_level(code) + 1, next);
  ls->step = nc;
  return nc;
}

// One block, launched after a level's kernels (step `code`): sums the
// blocks' partial totals (counts in cpart, degree sums in epart) plus what
// queue reservations took from ls->cnt, clears ls->cnt, decides the next
// step and posts the totals with its code.  The blocks of a level kernel
// finish together, so summing their totals here replaces same-address
// atomics, which serialize.  Thread 0 reads the loop state at the start
// (the earlier kernels' atomics are complete at this kernel's start), so
// its loads overlap the partial sums'.
__global__ __launch_bounds__(kT) void k_level_decide(const uint64_t* __restrict__ cpart,
                                                    int nc,
                                                    const uint64_t* __restrict__ epart,
                                                    int ne, LevelSync* ls, HostReport* hr,
                                                    unsigned long long seq, uint32_t code) {
  if (!step_is_next(ls, code)) return;
  __shared__ uint64_t red[kWavesPerBlock];
  LevelSync cur;
  if (threadIdx.x == 0) cur = *ls;
  uint64_t c = 0, d = 0;
  for (int i = threadIdx.x; i < nc; i += kT) c += cpart[i];
  for (int i = threadIdx.x; i < ne; i += kT) d += epart[i];
  c = block_sum(c, red);
  __syncthreads();  // thread 0 has read red
  d = block_sum(d, red);
  if (threadIdx.x == 0) {
    unsigned long long v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      v[i] = cur.cnt[i];
      ls->cnt[i] = 0;
    }
    v[0] += c;
    v[1] += d;
    v[2] = decide_next(cur, ls, code, (long long)v[0], (long long)v[1]);
    v[3] = 0;
    post_report(hr, v, seq);
  }
}

// Per-root setup of the level loop: first step, frontier, thresholds (the
// first frontier's degree sum is the root's row length, set on the device).
struct LoopInit {
  uint32_t first;  // code of the first step
  long long qlen, qedges, alpha_thr, beta_thr, exec_thr;
};

// Bits of vertices >= n_c are pre-set so they are never searched
// (the reference pre-visits vertices without edges, mpi/bfs.hpp:330-334).
// Per-root init in one launch: the bitmap (plus the root's bit) and, from
// thread 0, the root's pred entry, the one-vertex frontier and the level
// loop's state (cleared counters -- a run that failed inside a level leaves
// its reservations there -- and the first step).
__global__ __launch_bounds__(kT) void k_bfs_init(uint64_t* __restrict__ vis,
                                                int64_t words, int64_t n_c,
                                                int64_t root_new, int64_t* pred_new,
                                                int64_t entry, uint32_t* q, LevelSync* ls,
                                                LoopInit li,
                                                const int64_t* __restrict__ row_off) {
  for (int64_t w = (int64_t)blockIdx.x * kT + threadIdx.x; w < words;
       w += (int64_t)gridDim.x * kT) {
    const int64_t lo = w * 64;
    const int64_t valid = n_c - lo;
    uint64_t word = 0;
    if (valid < 64) word = valid <= 0 ? ~0ull : (~0ull << valid);
    if (root_new >= lo && root_new < lo + 64) word |= 1ull << (root_new - lo);
    vis[w] = word;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ls->cnt[i] = 0;
    ls->step = li.first;
    ls->growing = 1u;
    ls->qlen = li.qlen;
    ls->qedges = li.qedges;
    ls->prev_nq = 0;
    ls->alpha_thr = li.alpha_thr;
    ls->beta_thr = li.beta_thr;
    ls->exec_thr = li.exec_thr;
    ls->snap_tag = 0;
    if (root_new >= 0 && root_new < n_c) {
      pred_new[root_new] = entry;
      q[0] = (uint32_t)root_new;
      ls->qedges = row_off[root_new + 1] - row_off[root_new];
    }
  }
}

// Streams that no kernel of the BFS reads back soon (pred entries, the
// first-neighbour pairs, the unvisited lists) bypass the L2 with
// non-temporal accesses, which leaves it to the visited bitmap's probes.
__device__ inline void put_pred(int64_t* p, uint64_t entry) {
  __builtin_nontemporal_store((int64_t)entry, p);
}
__device__ inline uint2 get_pair(const uint2* p) {
  const unsigned long long x =
      __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(p));
  return make_uint2((uint32_t)x, (uint32_t)(x >> 32));
}

// Filter before the visited-bit atomicOr (which alone decides the winner).
__device__ inline unsigned long long probe_vis(unsigned long long* p) {
  return __atomic_load_n(p, __ATOMIC_RELAXED);
}

// largest i in [lo, hi] with a[i] <= e   (a non-decreasing, a[lo] <= e)
__device__ inline int64_t owner_search(const uint64_t* a, int64_t lo,
                                       int64_t hi, uint64_t e) {
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (a[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Top-down preparation of a frontier of more than kTdWin vertices, in two
// launches: dsum[i] = exclusive prefix sum of the frontier's degrees
// (dsum[qlen] = total) and tile_owner[t] = frontier index whose edge range
// holds edge t * kTdTile (the last vertex also writes the entry after the
// last tile).  Block b owns entries [b * kPrepTile, (b + 1) * kPrepTile),
// thread t of it kPrepItems consecutive ones.
constexpr int kPrepItems = 8;
constexpr int kPrepTile = kT * kPrepItems;

// The prep kernels run on a fixed grid (the frontier's size is on the
// device when they are queued): block b takes blocks b, b + grid, ... of
// the frontier.  Both return at once for a frontier of up to kTdWin vertices.
constexpr int kPrepGrid = 256;

// Block-local prefix sums into dsum, block totals into bsum; the last block
// to finish (atomic ticket, which it resets) scans the totals into boff.
__global__ __launch_bounds__(kT) void k_td_prep_sums(
    const uint32_t* __restrict__ q, const LevelSync* ls, uint32_t code,
    const int64_t* __restrict__ row_off, uint64_t* __restrict__ dsum,
    uint64_t* __restrict__ bsum, uint64_t* __restrict__ boff, unsigned int* ticket) {
  if (!step_is_next(ls, code)) return;
  const int64_t qlen = ls->qlen;
  if (qlen <= kTdWin) return;
  __shared__ uint64_t wsum[kWavesPerBlock];
  __shared__ bool last;
  const int t = threadIdx.x;
  const int64_t nb = ceil_div(qlen, (int64_t)kPrepTile);
  for (int64_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
    const int64_t i0 = blk * kPrepTile + (int64_t)t * kPrepItems;
    uint32_t u[kPrepItems];
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) u[j] = i0 + j < qlen ? q[i0 + j] : 0u;
    uint64_t d[kPrepItems];
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j)
      d[j] = i0 + j < qlen ? (uint64_t)(row_off[u[j] + 1] - row_off[u[j]]) : 0u;
    uint64_t sum = 0;
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) sum += d[j];
    uint64_t tot;
    uint64_t a = block_excl_scan(sum, wsum, &tot);
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) {
      if (i0 + j < qlen) dsum[i0 + j] = a;
      a += d[j];
    }
    if (t == 0) bsum[blk] = tot;
  }
  if (t == 0) {
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  uint64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kT) {
    const uint64_t x = b0 + t < nb ? __atomic_load_n(&bsum[b0 + t], __ATOMIC_RELAXED) : 0;
    uint64_t ctot;
    const uint64_t ex = block_excl_scan(x, wsum, &ctot);
    if (b0 + t < nb) boff[b0 + t] = carry + ex;
    carry += ctot;
  }
  if (t == 0) {
    dsum[qlen] = carry;
    *ticket = 0u;
  }
}

// Adds the block offsets and writes the tile owners.
__global__ __launch_bounds__(kT) void k_td_prep_tiles(
    const LevelSync* ls, uint32_t code, uint64_t* __restrict__ dsum,
    const uint64_t* __restrict__ bsum, const uint64_t* __restrict__ boff,
    uint32_t* __restrict__ tile_owner) {
  if (!step_is_next(ls, code)) return;
  const int64_t qlen = ls->qlen;
  if (qlen <= kTdWin) return;
  const int t = threadIdx.x;
  const int64_t nb = ceil_div(qlen, (int64_t)kPrepTile);
  for (int64_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
    const int64_t i0 = blk * kPrepTile + (int64_t)t * kPrepItems;
    const int64_t bend = min((blk + 1) * (int64_t)kPrepTile, qlen);
    // local prefix sums of this thread's entries and of the entry after them
    // (the next thread's first, or the block total), read before any write
    uint64_t loc[kPrepItems + 1];
#pragma unroll
    for (int j = 0; j <= kPrepItems; ++j) {
      const int64_t i = i0 + j;
      loc[j] = i < bend ? dsum[i] : bsum[blk];
    }
    const uint64_t off = boff[blk];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPrepItems; ++j) {
      const int64_t i = i0 + j;
      if (i >= bend) break;
      const uint64_t a = off + loc[j], b = off + loc[j + 1];
      dsum[i] = a;
      if (b <= a) continue;
      const uint64_t t0 = (a + kTdTile - 1) / kTdTile, t1 = (b - 1) / kTdTile;
      for (uint64_t tt = t0; tt <= t1; ++tt) tile_owner[tt] = (uint32_t)i;
      if (i == qlen - 1) tile_owner[t1 + 1] = (uint32_t)i;
    }
  }
}

// Frontiers of at most kTdWin vertices need no preparation kernels: every
// block stages the whole frontier (prefix sum of degrees, row base, original
// ID) in LDS itself.
// The grid is fixed (kMaxGrid); blocks past the frontier's degree sum (an
// upper bound of its row entries) write a zero total and return.
__global__ __launch_bounds__(kT) void k_td_expand(
    const uint32_t* __restrict__ q, uint32_t code,
    const uint64_t* __restrict__ dsum, const uint32_t* __restrict__ tile_owner,
    const int64_t* __restrict__ row_off,
    const uint32_t* __restrict__ col, const uint32_t* __restrict__ n2o,
    const uint32_t* __restrict__ deg_new, unsigned long long* vis,
    int64_t* __restrict__ pred_new, uint32_t* __restrict__ nq, LevelSync* ls,
    uint64_t* __restrict__ epart, uint64_t level_bits) {
  if (!step_is_next(ls, code)) return;
  if ((uint64_t)blockIdx.x * kTdTile >= (uint64_t)ls->qedges) {
    if (threadIdx.x == 0) epart[blockIdx.x] = 0;
    return;
  }
  const int64_t qlen = ls->qlen;
  // the tile's owners, staged once per tile: prefix sum, col offset of edge
  // 0 of the owner's range (row_off[u] - dsum[i]) and original ID.  On heavy
  // levels a tile has one or two owners (hubs) shared by every block at
  // once; reading them per lane from memory would pile thousands of
  // requests onto a few cache lines
  __shared__ uint64_t win[kTdWin];
  __shared__ int64_t wrb[kTdWin];
  __shared__ uint32_t wpu[kTdWin];
  // winners of the block's tiles, flushed to the next queue with one global
  // reservation once half a tile's worth has gathered (and after the last
  // tile): a reservation per tile serialized ~12 ns each on one counter,
  // which on levels of 10^8 edges (5*10^4 tiles) was most of their time
  __shared__ uint32_t lq[kTdTile + kTdTile / 2];
  __shared__ uint64_t red[kWavesPerBlock];
  __shared__ int64_t s_lo, s_hi;
  __shared__ uint32_t lq_n;
  __shared__ unsigned long long lq_base;

  const int t = threadIdx.x;
  const bool small = qlen <= kTdWin;
  uint64_t E;
  if (small) {
    // lane t stages frontier entries 2t and 2t+1
    const int64_t i0 = 2 * t, i1 = 2 * t + 1;
    uint32_t u0 = 0, u1 = 0;
    int64_t r0 = 0, r1 = 0;
    uint64_t d0 = 0, d1 = 0;
    if (i0 < qlen) {
      u0 = q[i0];
      r0 = row_off[u0];
      d0 = (uint64_t)(row_off[u0 + 1] - r0);
    }
    if (i1 < qlen) {
      u1 = q[i1];
      r1 = row_off[u1];
      d1 = (uint64_t)(row_off[u1 + 1] - r1);
    }
    const uint64_t ex = block_excl_scan(d0 + d1, red, &E);
    if (i0 < qlen) {
      win[i0] = ex;
      wrb[i0] = r0 - (int64_t)ex;
      wpu[i0] = n2o[u0];
    }
    if (i1 < qlen) {
      win[i1] = ex + d0;
      wrb[i1] = r1 - (int64_t)(ex + d0);
      wpu[i1] = n2o[u1];
    }
    __syncthreads();
  } else {
    E = dsum[qlen];
  }
  uint64_t my_edges = 0;
  if (t == 0) lq_n = 0;
  for (uint64_t e0 = (uint64_t)blockIdx.x * kTdTile; e0 < E;
       e0 += (uint64_t)gridDim.x * kTdTile) {
    const uint64_t e1 = (e0 + kTdTile < E) ? e0 + kTdTile : E;
    if (t == 0) {
      // owners of the tile's first edge and of the next tile's first edge
      // (>= the owner of its last edge)
      const uint64_t tile = e0 / kTdTile;
      s_lo = small ? 0 : tile_owner[tile];
      s_hi = small ? qlen - 1 : min((int64_t)tile_owner[tile + 1], qlen - 1);
    }
    __syncthreads();
    const int64_t lo = s_lo, hi = s_hi;
    const int64_t wn = hi - lo + 1;
    const bool use_win = wn <= kTdWin;
    if (use_win && !small)
      for (int64_t j = t; j < wn; j += kT) {
        const uint64_t d = dsum[lo + j];
        const uint32_t u = q[lo + j];
        win[j] = d;
        wrb[j] = row_off[u] - (int64_t)d;
        wpu[j] = n2o[u];
      }
    __syncthreads();
    // item j of lane t is edge e0 + j*kT + t (coalesced col reads); the loads
    // of all items are issued phase by phase so their latencies overlap
    uint32_t v[kTdItems], pu[kTdItems];
    bool act[kTdItems];
    {
      int64_t kk[kTdItems];
#pragma unroll
      for (int j = 0; j < kTdItems; ++j) {
        const uint64_t e = e0 + (uint64_t)j * kT + t;
        act[j] = e < e1;
        kk[j] = 0;
        pu[j] = 0;
        if (act[j]) {
          if (use_win) {
            const int64_t li = owner_search(win, 0, wn - 1, e);
            kk[j] = wrb[li] + (int64_t)e;
            pu[j] = wpu[li];
          } else {
            const int64_t i = owner_search(dsum, lo, hi, e);
            const uint32_t u = q[i];
            kk[j] = row_off[u] + (int64_t)(e - dsum[i]);
            pu[j] = n2o[u];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kTdItems; ++j) v[j] = act[j] ? col[kk[j]] : 0u;
    }
    uint64_t w[kTdItems];
#pragma unroll
    for (int j = 0; j < kTdItems; ++j)
      w[j] = act[j] ? probe_vis(vis + (v[j] >> 6)) : ~0ull;
    uint32_t nwon = 0;
    bool won[kTdItems];
#pragma unroll
    for (int j = 0; j < kTdItems; ++j) {
      const unsigned long long bit = 1ull << (v[j] & 63);
      won[j] = false;
      if (act[j] && !(w[j] & bit)) won[j] = !(atomicOr(vis + (v[j] >> 6), bit) & bit);
      nwon += won[j];
    }
#pragma unroll
    for (int j = 0; j < kTdItems; ++j)
      if (won[j]) {
        put_pred(pred_new + v[j], (uint64_t)pu[j] | level_bits);
        my_edges += deg_new[v[j]];
      }
    // append winners to the block's LDS queue: wave prefix + one LDS atomic
    {
      const int lane = lane_id();
      uint32_t incl = nwon;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += y;
      }
      const uint32_t wtot = __shfl(incl, kWave - 1, kWave);
      uint32_t wbase = 0;
      if (lane == kWave - 1 && wtot) wbase = atomicAdd(&lq_n, wtot);
      wbase = __shfl(wbase, kWave - 1, kWave);
      uint32_t pos = wbase + incl - nwon;
#pragma unroll
      for (int j = 0; j < kTdItems; ++j)
        if (won[j]) lq[pos++] = v[j];
    }
    __syncthreads();
    // below kTdTile / 2 left, the next tile's winners (at most kTdTile) fit
    const uint32_t cnt = lq_n;
    const bool last = e0 + (uint64_t)gridDim.x * kTdTile >= E;
    if (cnt >= kTdTile / 2 || (last && cnt)) {
      if (t == 0) lq_base = atomicAdd(&ls->cnt[0], (unsigned long long)cnt);
      __syncthreads();
      for (uint32_t j = t; j < cnt; j += kT) nq[lq_base + j] = lq[j];
      __syncthreads();
      if (t == 0) lq_n = 0;
    }
    __syncthreads();
  }
  const uint64_t s = block_sum(my_edges, red);
  if (t == 0) epart[blockIdx.x] = s;
}

// Bottom-up levels.  A wave owns a chunk of 64 bitmap words (4096 vertices)
// and, for it, a segment of 4096 entries in the unvisited lists: after a
// bottom-up level the segment's first unv_cnt[chunk] entries are the chunk's
// still-unvisited vertices.  The first bottom-up level after top-down ones
// sweeps the bitmap (k_bu_sweep); later ones walk the lists (k_bu_list),
// which shrink level by level.  A vertex first probes its highest-degree
// neighbour (the dense `first` pairs also give that neighbour's original ID
// for pred): nearly every vertex found in a bottom-up level is found there.
// Only vertices whose row has more entries (the `multi` bitmap) scan the
// rest of it; the sweep defers them to the back of the chunk's segment and
// scans them together at the end of the chunk.  A visited neighbour of a
// vertex unvisited at the start of a level lies on the previous level, so
// the visited bitmap serves as the frontier (bottom_up_search_bitmap,
// mpi/bfs.hpp:2828-2978; the reference also keeps unvisited vertices as
// lists, swap_visited_memory and bottom_up_search_list, mpi/bfs.hpp:1747-1779
// and 2979-3005).  The parent is the first visited neighbour in row order.
constexpr int kBuIlp = 8;   // vertices in flight per lane
constexpr int kBuBatch = kWave * kBuIlp;
constexpr int kChunkVerts = kWave * 64;
// The visited bits of the lowest new IDs (the highest degrees) sit in LDS:
// most first neighbours are such hubs, so most probes never leave the CU.
constexpr int kHubWords = 4096;
// The bottom-up kernels' block size: all waves of a block share its hub copy.
constexpr int kBuT = 256;
constexpr int kBuWaves = kBuT / kWave;

// Visited bit of u: from the block's LDS copy below hub_verts, else memory.
struct VisProbe {
  const uint64_t* hub;
  const uint64_t* vis;
  uint32_t hub_verts;
  __device__ bool operator()(uint32_t u) const {
    const uint64_t w = u < hub_verts ? hub[u >> 6] : vis[u >> 6];
    return (w >> (u & 63)) & 1ull;
  }
};

// Copies the first min(words, kHubWords) words of vis into LDS (all threads).
__device__ inline VisProbe stage_hubs(uint64_t* hub, const uint64_t* vis, int64_t words) {
  const int hw = (int)(words < kHubWords ? words : kHubWords);
  for (int i = threadIdx.x; i < hw; i += kBuT) hub[i] = vis[i];
  __syncthreads();
  return VisProbe{hub, vis, (uint32_t)hw * 64u};
}

__device__ inline uint64_t readlane64(uint64_t x, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Rows with more than max_rest entries left are not scanned: long[j] is set
// for them instead.
// With `count`, *scan_bytes grows by the bytes the scans need: the row's
// offsets (16 B), each entry read up to the first visited one (4 B) and the
// parent's original ID (4 B).
__device__ inline void bu_rows(const uint32_t (&v)[kBuIlp], const bool (&scan)[kBuIlp],
                               bool (&hit)[kBuIlp], bool (&lng)[kBuIlp],
                               int64_t max_rest, const VisProbe& visited,
                               const int64_t* __restrict__ row_off,
                               const uint32_t* __restrict__ col,
                               const uint32_t* __restrict__ n2o,
                               int64_t* __restrict__ pred_new, uint64_t level_bits,
                               bool count, uint64_t* scan_bytes) {
  int64_t k[kBuIlp], k1[kBuIlp];
#pragma unroll
  for (int j = 0; j < kBuIlp; ++j) {
    k[j] = scan[j] ? row_off[v[j]] + 1 : 0;
    k1[j] = scan[j] ? row_off[v[j] + 1] : 0;
    lng[j] = k1[j] - k[j] > max_rest;
    if (lng[j]) k1[j] = k[j];
    if (count && scan[j]) *scan_bytes += 16;
  }
  uint32_t par[kBuIlp];
#pragma unroll
  for (int j = 0; j < kBuIlp; ++j) {
    hit[j] = false;
    par[j] = 0;
    while (k[j] < k1[j]) {
      uint32_t u[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) u[t] = (k[j] + t < k1[j]) ? col[k[j] + t] : 0u;
      bool h[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) h[t] = (k[j] + t < k1[j]) && visited(u[t]);
      int used = (int)min<int64_t>(4, k1[j] - k[j]);
#pragma unroll
      for (int t = 3; t >= 0; --t)
        if (h[t]) {
          hit[j] = true;
          par[j] = u[t];
          used = t + 1;
        }
      if (count) *scan_bytes += 4u * (uint64_t)used + (hit[j] ? 4u : 0u);
      k[j] = hit[j] ? k1[j] : k[j] + 4;
    }
  }
  uint32_t pp[kBuIlp];
#pragma unroll
  for (int j = 0; j < kBuIlp; ++j) pp[j] = hit[j] ? n2o[par[j]] : 0u;
#pragma unroll
  for (int j = 0; j < kBuIlp; ++j)
    if (hit[j]) put_pred(pred_new + v[j], (uint64_t)pp[j] | level_bits);
}

// Appends the lanes' flagged vertices at useg[pos ...] (ascending) or, with
// down, at useg[pos - 1], useg[pos - 2], ...; returns how many.
__device__ inline uint32_t bu_append(const uint32_t (&v)[kBuIlp], const bool (&flag)[kBuIlp],
                                     uint32_t* __restrict__ useg, uint32_t pos, bool down) {
  const uint64_t below = lanemask_lt();
  uint32_t added = 0;
#pragma unroll
  for (int j = 0; j < kBuIlp; ++j) {
    const uint64_t m = __ballot(flag[j]);
    const uint32_t r = added + (uint32_t)__popcll(m & below);
    if (flag[j]) __builtin_nontemporal_store(v[j], useg + (down ? pos - 1 - r : pos + r));
    added += (uint32_t)__popcll(m);
  }
  return added;
}

__device__ inline bool multi_bit(const uint64_t* multi, uint32_t v) {
  return (multi[v >> 6] >> (v & 63)) & 1ull;
}

// Rows with more than kBuLongRow entries left after a missed first
// neighbour are not scanned by k_bu_sweep's lanes: they are listed for
// k_bu_long, which gives each one a whole wave (one such row would otherwise
// hold its wave far longer than the rest).  k_bu_long closes the level.
constexpr int64_t kBuLongRow = 256;
constexpr int kBuStep = 8;  // row entries a lane probes per step (two uint4 loads)

__global__ __launch_bounds__(kBuT) void k_bu_sweep(
    const uint64_t* __restrict__ vis_old, uint64_t* __restrict__ vis_new,
    int64_t words, const int64_t* __restrict__ row_off,
    const uint32_t* __restrict__ col, const uint2* __restrict__ first,
    const uint64_t* __restrict__ multi, const uint32_t* __restrict__ n2o,
    int64_t* __restrict__ pred_new, uint32_t* __restrict__ unv,
    uint32_t* __restrict__ unv_cnt, uint32_t* __restrict__ longv, LevelSync* ls,
    uint32_t code, uint64_t* __restrict__ cpart, uint64_t level_bits,
    uint64_t* __restrict__ wacc) {
  if (!step_is_next(ls, code)) return;
  __shared__ unsigned long long fw[kBuWaves][kWave];
  __shared__ uint64_t red[kBuWaves];
  __shared__ uint64_t hub[kHubWords];
  const VisProbe visited = stage_hubs(hub, vis_old, words);
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int64_t nchunks = ceil_div(words, kWave);
  const bool count = wacc != nullptr;  // counting runs tally the row-scan bytes
  uint64_t nfound = 0, scan_b = 0;
  for (int64_t c = (int64_t)blockIdx.x * kBuWaves + wave; c < nchunks;
       c += (int64_t)gridDim.x * kBuWaves) {
    const int64_t w = c * kWave + lane;
    const uint64_t myword = w < words ? vis_old[w] : ~0ull;
    const uint64_t mymulti = w < words ? multi[w] : 0ull;
    uint32_t* useg = unv + c * kChunkVerts;
    const uint32_t vbase = (uint32_t)(c * kChunkVerts);
    uint64_t newbits = 0;  // lane l: bits found in word l of the chunk
    // final misses fill the segment from the front, deferred row scans from
    // the back; together they never exceed the chunk's unvisited vertices
    uint32_t wp = 0, dn = 0;
    // lane l takes bit l of words j0 .. j0+kBuIlp-1: coalesced `first`
    // reads.  The next batch's pairs are requested before this batch's
    // probes and stores, so their latency overlaps the work (vector memory
    // counters retire in order, so loads issued after a store would also
    // wait for it)
    uint2 fn[kBuIlp];
    bool okn[kBuIlp];
#pragma unroll
    for (int j = 0; j < kBuIlp; ++j) {
      okn[j] = !((readlane64(myword, j) >> lane) & 1ull);
      fn[j] = okn[j] ? get_pair(first + vbase + (uint32_t)(j * 64 + lane)) : make_uint2(0u, 0u);
    }
    for (int j0 = 0; j0 < kWave; j0 += kBuIlp) {
      uint32_t v[kBuIlp];
      uint2 f[kBuIlp];
      bool ok[kBuIlp], hit[kBuIlp], fin[kBuIlp], defer[kBuIlp];
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) {
        ok[j] = okn[j];
        f[j] = fn[j];
        v[j] = vbase + (uint32_t)((j0 + j) * 64 + lane);
      }
      if (j0 + kBuIlp < kWave) {
#pragma unroll
        for (int j = 0; j < kBuIlp; ++j) {
          okn[j] = !((readlane64(myword, j0 + kBuIlp + j) >> lane) & 1ull);
          fn[j] = okn[j] ? get_pair(first + v[j] + kBuIlp * 64) : make_uint2(0u, 0u);
        }
      }
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) hit[j] = ok[j] && visited(f[j].x);
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j)
        if (hit[j]) put_pred(pred_new + v[j], (uint64_t)f[j].y | level_bits);
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) {
        const uint64_t hm = __ballot(hit[j]);
        if (lane == j0 + j) newbits = hm;
        const bool more = (readlane64(mymulti, j0 + j) >> lane) & 1ull;
        fin[j] = ok[j] && !hit[j] && !more;
        defer[j] = ok[j] && !hit[j] && more;
      }
      // one store per slot: final misses to the front, deferred rows to the back
      {
        const uint64_t below = lanemask_lt();
        uint32_t addf = 0, addd = 0;
#pragma unroll
        for (int j = 0; j < kBuIlp; ++j) {
          const uint64_t mf = __ballot(fin[j]), md = __ballot(defer[j]);
          if (fin[j] || defer[j]) {
            const uint32_t pos =
                fin[j] ? wp + addf + (uint32_t)__popcll(mf & below)
                       : kChunkVerts - dn - 1 - (addd + (uint32_t)__popcll(md & below));
            __builtin_nontemporal_store(v[j], useg + pos);
          }
          addf += (uint32_t)__popcll(mf);
          addd += (uint32_t)__popcll(md);
        }
        wp += addf;
        dn += addd;
      }
    }
    // deferred rows: a lane takes the next one (lowest segment position
    // first) whenever it has none, and scans kBuStep entries per step, so a
    // wave's time follows its total work rather than its longest row per
    // batch.  The front grows only over entries already read (each row
    // finished was taken before).  Rows longer than kBuLongRow go to
    // k_bu_long.
    fw[wave][lane] = 0;
    wave_sync();
    {
      const uint32_t dbase = kChunkVerts - dn;
      uint32_t taken = 0;  // wave-uniform
      bool have = false;
      uint32_t dv = 0;
      int64_t dlo = 0, dk = 0, dk1 = 0;  // row part [dlo, dk1); next step at dk (16 B aligned)
      const uint64_t below = lanemask_lt();
      for (;;) {
        const uint64_t idle = __ballot(!have);
        const uint32_t idx = taken + (uint32_t)__popcll(idle & below);
        taken += (uint32_t)__popcll(idle);
        if (taken > dn) taken = dn;
        bool lng = false;
        if (!have && idx < dn) {
          dv = useg[dbase + idx];
          dlo = row_off[dv] + 1;
          dk1 = row_off[dv + 1];
          dk = dlo & ~(int64_t)3;
          have = true;
          lng = dk1 - dlo > kBuLongRow;
          // row offsets, and the deferred entry written and read back
          if (count) scan_b += 24;
        }
        if (__ballot(have) == 0) break;
        const uint64_t lm = __ballot(lng);
        if (lm) {
          unsigned long long base = 0;
          if (lane == 0) base = atomicAdd(&ls->cnt[2], (unsigned long long)__popcll(lm));
          base = __shfl(base, 0, kWave);
          if (lng) {
            longv[base + (uint64_t)__popcll(lm & below)] = dv;
            have = false;
          }
        }
        bool hit = false, miss = false;
        uint32_t par = 0;
        if (have) {
          // two aligned 16-byte loads per step (col carries a few padding
          // entries past m); entries outside [dlo, dk1) are not probed
          const uint4* c4 = reinterpret_cast<const uint4*>(col) + (dk >> 2);
          const uint4 a = c4[0];
          const uint4 b4 = dk + 4 < dk1 ? c4[1] : make_uint4(0u, 0u, 0u, 0u);
          const uint32_t u[kBuStep] = {a.x, a.y, a.z, a.w, b4.x, b4.y, b4.z, b4.w};
          bool h[kBuStep];
#pragma unroll
          for (int t = 0; t < kBuStep; ++t)
            h[t] = dk + t >= dlo && dk + t < dk1 && visited(u[t]);
          int stop = (int)min<int64_t>(kBuStep, dk1 - dk);
#pragma unroll
          for (int t = kBuStep - 1; t >= 0; --t)
            if (h[t]) {
              hit = true;
              par = u[t];
              stop = t + 1;
            }
          if (count) {
            // entries of the row read in this step, up to the first visited one
            const int start = dlo > dk ? (int)(dlo - dk) : 0;
            scan_b += 4u * (uint64_t)max(stop - start, 0) + (hit ? 4u : 0u);
          }
          dk += kBuStep;
          miss = !hit && dk >= dk1;
        }
        if (hit) {
          put_pred(pred_new + dv, (uint64_t)n2o[par] | level_bits);
          atomicOr(&fw[wave][(dv - vbase) >> 6], 1ull << (dv & 63));
        }
        const uint64_t mm = __ballot(miss);
        if (miss) useg[wp + (uint32_t)__popcll(mm & below)] = dv;
        wp += (uint32_t)__popcll(mm);
        if (hit || miss) have = false;
      }
    }
    wave_sync();
    newbits |= fw[wave][lane];
    if (w < words) vis_new[w] = myword | newbits;
    nfound += (uint64_t)__popcll(newbits);
    if (lane == 0) unv_cnt[c] = wp;
    // the entries of the vertices found, and the chunk's unvisited list
    if (count) scan_b += 8u * (uint64_t)__popcll(newbits) + (lane == 0 ? 4u * (uint64_t)wp : 0u);
    wave_sync();
  }
  const uint64_t s = block_sum<kBuWaves>(nfound, red);
  if (threadIdx.x == 0) cpart[blockIdx.x] = s;
  if (count) {
    __syncthreads();  // thread 0 has read red
    const uint64_t sb = block_sum<kBuWaves>(scan_b, red);
    if (threadIdx.x == 0) wacc[blockIdx.x] += sb;  // the block's own slot
  }
}

// The rows k_bu_sweep listed, a wave each: 256 entries per step, first
// visited neighbour in row order; a vertex without one joins its chunk's
// unvisited list.
__global__ __launch_bounds__(kBuT) void k_bu_long(
    const uint64_t* __restrict__ vis_old, unsigned long long* vis_new, int64_t words,
    const int64_t* __restrict__ row_off, const uint32_t* __restrict__ col,
    const uint32_t* __restrict__ n2o, int64_t* __restrict__ pred_new,
    uint32_t* __restrict__ unv, uint32_t* unv_cnt,
    const uint32_t* __restrict__ longv, const LevelSync* ls, uint32_t code,
    uint64_t* __restrict__ cpart, uint64_t level_bits, uint64_t* __restrict__ wacc) {
  if (!step_is_next(ls, code)) return;
  __shared__ uint64_t red[kBuWaves];
  __shared__ uint64_t hub[kHubWords];
  const VisProbe visited = stage_hubs(hub, vis_old, words);
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int64_t n_long = (int64_t)ls->cnt[2];
  const bool count = wacc != nullptr;
  uint64_t nfound = 0, scan_b = 0;
  for (int64_t it = (int64_t)blockIdx.x * kBuWaves + wave; it < n_long;
       it += (int64_t)gridDim.x * kBuWaves) {
    const uint32_t v = longv[it];
    const int64_t k0 = row_off[v] + 1;
    int64_t k = k0;
    const int64_t k1 = row_off[v + 1];
    bool found = false;
    uint32_t par = 0;
    int64_t at = k1;  // entry holding the first visited neighbour
    while (!found && k < k1) {
      uint32_t u[4];
      bool h[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t e = k + i * kWave + lane;
        u[i] = e < k1 ? col[e] : 0u;
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) h[i] = k + i * kWave + lane < k1 && visited(u[i]);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint64_t m = __ballot(h[i]);
        if (m && !found) {
          found = true;
          par = __shfl(u[i], __ffsll((long long)m) - 1, kWave);
          at = k + i * kWave + (__ffsll((long long)m) - 1);
        }
      }
      k += 4 * kWave;
    }
    // row offsets, the listed entry, entries up to the first visited one,
    // then the parent's original ID or the unvisited-list entry
    if (count && lane == 0)
      scan_b += 20 + 4 * (uint64_t)((found ? at + 1 : k1) - k0) + (found ? 12 : 4);
    if (lane == 0) {
      if (found) {
        put_pred(pred_new + v, (uint64_t)n2o[par] | level_bits);
        atomicOr(vis_new + (v >> 6), 1ull << (v & 63));
        ++nfound;
      } else {
        const uint32_t c = v / (uint32_t)kChunkVerts;
        const uint32_t pos = atomicAdd(unv_cnt + c, 1u);
        unv[(int64_t)c * kChunkVerts + pos] = v;
      }
    }
  }
  const uint64_t s = block_sum<kBuWaves>(nfound, red);
  if (threadIdx.x == 0) cpart[blockIdx.x] = s;
  if (count) {
    __syncthreads();
    const uint64_t sb = block_sum<kBuWaves>(scan_b, red);
    if (threadIdx.x == 0) wacc[blockIdx.x] += sb;
  }
}

// Four waves per SIMD (128 VGPRs, a 12-byte spill) rather than the three its
// registers would give: its LDS allows four, and the list levels are
// latency-bound.
__global__ __launch_bounds__(kBuT) __attribute__((amdgpu_waves_per_eu(4))) void k_bu_list(
    const uint64_t* __restrict__ vis_old, uint64_t* __restrict__ vis_new,
    int64_t words, const int64_t* __restrict__ row_off,
    const uint32_t* __restrict__ col, const uint2* __restrict__ first,
    const uint64_t* __restrict__ multi, const uint32_t* __restrict__ n2o,
    int64_t* __restrict__ pred_new, uint32_t* __restrict__ unv,
    uint32_t* __restrict__ unv_cnt, const LevelSync* ls, uint32_t code,
    uint64_t* __restrict__ cpart, uint64_t level_bits, uint64_t* __restrict__ wacc) {
  if (!step_is_next(ls, code)) return;
  __shared__ unsigned long long fw[kBuWaves][kWave];
  __shared__ uint64_t red[kBuWaves];
  __shared__ uint64_t hub[kHubWords];
  const VisProbe visited = stage_hubs(hub, vis_old, words);
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int64_t nchunks = ceil_div(words, kWave);
  const bool count = wacc != nullptr;
  uint64_t nfound = 0, scan_b = 0;
  for (int64_t c = (int64_t)blockIdx.x * kBuWaves + wave; c < nchunks;
       c += (int64_t)gridDim.x * kBuWaves) {
    const int64_t w = c * kWave + lane;
    const uint64_t myword = w < words ? vis_old[w] : ~0ull;
    const uint32_t n = unv_cnt[c];
    uint32_t* useg = unv + c * kChunkVerts;
    const uint32_t vbase = (uint32_t)(c * kChunkVerts);
    fw[wave][lane] = 0;
    wave_sync();
    // compaction in place: misses of a batch land below the next batch, so
    // the write position never passes an entry not yet read.  The next
    // batch's entries and first-neighbour pairs are requested before this
    // batch's probes and stores (as in the sweep).
    uint32_t wp = 0;
    uint32_t vn[kBuIlp];
    uint2 fn[kBuIlp];
#pragma unroll
    for (int j = 0; j < kBuIlp; ++j) {
      const uint32_t i = (uint32_t)(j * kWave + lane);
      vn[j] = i < n ? useg[i] : vbase;
    }
#pragma unroll
    for (int j = 0; j < kBuIlp; ++j)
      fn[j] = (uint32_t)(j * kWave + lane) < n ? get_pair(first + vn[j]) : make_uint2(0u, 0u);
    for (uint32_t e0 = 0; e0 < n; e0 += kBuBatch) {
      uint32_t v[kBuIlp];
      uint2 f[kBuIlp];
      bool ok[kBuIlp], hit[kBuIlp], scan[kBuIlp], hit2[kBuIlp], miss[kBuIlp];
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) {
        ok[j] = e0 + (uint32_t)(j * kWave + lane) < n;
        v[j] = vn[j];
        f[j] = fn[j];
      }
      if (e0 + kBuBatch < n) {
#pragma unroll
        for (int j = 0; j < kBuIlp; ++j) {
          const uint32_t i = e0 + kBuBatch + (uint32_t)(j * kWave + lane);
          vn[j] = i < n ? useg[i] : vbase;
        }
#pragma unroll
        for (int j = 0; j < kBuIlp; ++j)
          fn[j] = e0 + kBuBatch + (uint32_t)(j * kWave + lane) < n ? get_pair(first + vn[j])
                                                                   : make_uint2(0u, 0u);
      }
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) hit[j] = ok[j] && visited(f[j].x);
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j)
        if (hit[j]) put_pred(pred_new + v[j], (uint64_t)f[j].y | level_bits);
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) scan[j] = ok[j] && !hit[j] && multi_bit(multi, v[j]);
      bool lng[kBuIlp];
      bu_rows(v, scan, hit2, lng, INT64_MAX, visited, row_off, col, n2o, pred_new, level_bits,
              count, &scan_b);
#pragma unroll
      for (int j = 0; j < kBuIlp; ++j) {
        hit[j] = hit[j] || hit2[j];
        if (hit[j]) atomicOr(&fw[wave][(v[j] - vbase) >> 6], 1ull << (v[j] & 63));
        miss[j] = ok[j] && !hit[j];
      }
      wp += bu_append(v, miss, useg, wp, false);
    }
    wave_sync();
    const uint64_t fbits = fw[wave][lane];
    if (w < words) vis_new[w] = myword | fbits;
    nfound += (uint64_t)__popcll(fbits);
    if (lane == 0) unv_cnt[c] = wp;
    if (count) scan_b += 8u * (uint64_t)__popcll(fbits) + (lane == 0 ? 4u * (uint64_t)wp : 0u);
    wave_sync();
  }
  const uint64_t s = block_sum<kBuWaves>(nfound, red);
  if (threadIdx.x == 0) cpart[blockIdx.x] = s;
  if (count) {
    __syncthreads();
    const uint64_t sb = block_sum<kBuWaves>(scan_b, red);
    if (threadIdx.x == 0) wacc[blockIdx.x] += sb;
  }
}

// queue <- bits set in cur but not in prev; reports (count, degree sum).
// Two launches over one partition of the words: block b owns a contiguous
// range and each of its waves a contiguous quarter of it.  k_queue_count
// writes every wave's count (and, when asked, the snapshot copy of cur for
// the pred gather); k_queue_write takes its offset from the counts of the
// waves before it, so no block waits on a shared reservation, and each wave
// compacts its own words with ballots: lanes store the queue and load the
// degrees coalesced, with up to four words' degree loads in flight.
constexpr int kQueueGrid = 1024;

__device__ inline void queue_wave_span(int64_t words, int64_t* a, int64_t* e) {
  const int64_t per_block = ceil_div(words, (int64_t)gridDim.x);
  const int64_t b0 = min((int64_t)blockIdx.x * per_block, words);
  const int64_t b1 = min(b0 + per_block, words);
  const int64_t per_wave = ceil_div(b1 - b0, (int64_t)kWavesPerBlock);
  *a = min(b0 + (int64_t)(threadIdx.x / kWave) * per_wave, b1);
  *e = min(*a + per_wave, b1);
}

__global__ __launch_bounds__(kT) void k_queue_count(const uint64_t* __restrict__ cur,
                                                   const uint64_t* __restrict__ prev,
                                                   int64_t words, uint32_t* __restrict__ wcnt,
                                                   uint64_t* __restrict__ snap,
                                                   LevelSync* ls, uint32_t code) {
  if (!step_is_next(ls, code)) return;
  // the gather queued for this conversion checks the tag
  if (snap && blockIdx.x == 0 && threadIdx.x == 0) ls->snap_tag = code;
  int64_t a, e;
  queue_wave_span(words, &a, &e);
  uint64_t c = 0;
  for (int64_t w = a + lane_id(); w < e; w += kWave) {
    const uint64_t x = cur[w];
    c += (uint64_t)__popcll(x & ~prev[w]);
    if (snap) snap[w] = x;
  }
  c = wave_sum_u64(c);
  if (lane_id() == 0) wcnt[(int64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave] = (uint32_t)c;
}

__global__ __launch_bounds__(kT) void k_queue_write(const uint64_t* __restrict__ cur,
                                                   const uint64_t* __restrict__ prev,
                                                   int64_t words,
                                                   const uint32_t* __restrict__ deg_new,
                                                   const uint32_t* __restrict__ wcnt,
                                                   uint32_t* __restrict__ out,
                                                   uint64_t* __restrict__ bedges,
                                                   const LevelSync* ls, uint32_t code) {
  if (!step_is_next(ls, code)) return;
  __shared__ uint64_t red[kWavesPerBlock];
  const int t = threadIdx.x, lane = lane_id(), wave = t / kWave;
  const int64_t first = (int64_t)blockIdx.x * kWavesPerBlock;
  uint64_t before = 0;
  for (int64_t i = t; i < first; i += kT) before += wcnt[i];
  before = wave_sum_u64(before);
  if (lane == 0) red[wave] = before;
  __syncthreads();
  uint64_t off = 0;
#pragma unroll
  for (int w = 0; w < kWavesPerBlock; ++w) off += red[w];
  for (int w = 0; w < wave; ++w) off += wcnt[first + w];
  int64_t a, e;
  queue_wave_span(words, &a, &e);
  const uint64_t below = (1ull << lane) - 1ull;
  uint64_t edges = 0;
  for (int64_t g = a; g < e; g += kWave) {
    const int64_t w = g + lane;
    const uint64_t mine = w < e ? (cur[w] & ~prev[w]) : 0ull;
    uint64_t nz = __ballot(mine != 0ull);
    while (nz) {
      uint32_t d[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        d[u] = 0;
        if (nz) {
          const int j = __ffsll((unsigned long long)nz) - 1;
          nz &= nz - 1;
          const uint64_t x = readlane64(mine, j);
          if ((x >> lane) & 1ull) {
            const uint32_t v = (uint32_t)((g + j) * 64 + lane);
            out[off + (uint64_t)__popcll(x & below)] = v;
            d[u] = deg_new[v];
          }
          off += (uint64_t)__popcll(x);
        }
      }
      edges += (uint64_t)d[0] + d[1] + d[2] + d[3];
    }
  }
  __syncthreads();  // red is reused
  const uint64_t s = block_sum(edges, red);
  if (t == 0) bedges[blockIdx.x] = s;
}

// One block, closing a step whose frontier came from the bitmap: sums the
// conversion's totals (count, degree sum) -- all blocks of k_queue_write
// finish together, so here rather than by atomics on one address, which
// serialize -- and clears ls->cnt (a top-down level run as a sweep leaves
// the long-row count there).  After a top-down level run as a sweep the
// totals are the level's and the direction rule decides the next step; after
// the conversion at the switch to top-down they are the first top-down
// frontier's, which runs next (as a sweep if its degree sum passes the
// execution threshold).
__global__ __launch_bounds__(kT) void k_queue_decide(const uint32_t* __restrict__ wcnt,
                                                    const uint64_t* __restrict__ bedges,
                                                    int blocks, LevelSync* ls, HostReport* hr,
                                                    unsigned long long seq, uint32_t code) {
  if (!step_is_next(ls, code)) return;
  __shared__ uint64_t red[kWavesPerBlock];
  LevelSync cur;
  if (threadIdx.x == 0) cur = *ls;
  uint64_t c = 0, d = 0;
  for (int i = threadIdx.x; i < blocks * kWavesPerBlock; i += kT) c += wcnt[i];
  for (int i = threadIdx.x; i < blocks; i += kT) d += bedges[i];
  c = block_sum(c, red);
  __syncthreads();  // thread 0 has read red
  d = block_sum(d, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ls->cnt[i] = 0;
    uint32_t next;
    if (step_mode(code) == kModeConv) {
      ls->qlen = (long long)c;
      ls->qedges = (long long)d;
      next = step_code(step_level(code), td_step_mode((long long)c, (long long)d, cur.exec_thr));
      ls->step = next;
    } else {
      next = decide_next(cur, ls, code, (long long)c, (long long)d);
    }
    const unsigned long long v[4] = {c, d, next, 0ull};
    post_report(hr, v, seq);
  }
}

// pred[o] for every original ID: the level-written entry of its new ID if
// that vertex was reached, the root's own entry, else -1.  With CoreBFS a
// forest vertex not reached through the core gets its forest parent with
// depth 0xFFFF (write_tree_parents, mpi/corebfs_adaptor.hpp:285-301).
// The GPU deals blocks to its 8 XCDs in turn (block b runs on XCD b % 8),
// and each XCD has its own L2.  The gathers read pred_new at new IDs that,
// within one degree class, rise with the original ID (the relabel breaks ties
// by original ID): neighbouring items share those cache lines.  So each XCD
// takes one contiguous eighth of the items and its blocks stride through
// that eighth together; the grid is a multiple of kXcds.
constexpr int kXcds = 8;
struct XcdSpan {
  int64_t lo, hi, rank, nblk;  // this XCD's items [lo, hi); block rank among its nblk
};
__device__ inline XcdSpan xcd_span(int64_t n) {
  const int64_t per = ceil_div(n, (int64_t)kXcds);
  XcdSpan x;
  x.lo = min((int64_t)(blockIdx.x % kXcds) * per, n);
  x.hi = min(x.lo + per, n);
  x.rank = blockIdx.x / kXcds;
  x.nblk = gridDim.x / kXcds;
  return x;
}

// Level entries carry the run's epoch in bits 32-47 (a parent's original ID
// fits 32 bits): an entry is this run's iff its epoch matches, so the
// gathers need no visited-bitmap probe, and pred gets the entry with the
// epoch cleared.
__device__ inline int64_t run_entry(int64_t e, uint32_t epoch) {
  return (((uint64_t)e >> 32) & 0xFFFFull) == epoch ? (int64_t)((uint64_t)e & ~kEpochMask)
                                                     : (int64_t)-1;
}

// A gather queued on a guess that a conversion runs (snap_code != 0) returns
// at once unless that conversion wrote the bitmap snapshot.
__device__ inline bool snapshot_is(const LevelSync* ls, uint32_t snap_code) {
  return snap_code == 0 || __builtin_amdgcn_readfirstlane(ls->snap_tag) == snap_code;
}

__global__ __launch_bounds__(kT) void k_pred_gather(const uint32_t* __restrict__ o2n,
                                                   uint32_t epoch,
                                                   const int64_t* __restrict__ pred_new,
                                                   int64_t n, int64_t n_c, int64_t root,
                                                   const uint32_t* __restrict__ tree_parent,
                                                   const uint64_t* __restrict__ has_parent,
                                                   int64_t* __restrict__ pred,
                                                   const LevelSync* ls, uint32_t snap_code) {
  if (!snapshot_is(ls, snap_code)) return;
  const XcdSpan xs = xcd_span(n);
  for (int64_t o = xs.lo + xs.rank * kT + threadIdx.x; o < xs.hi; o += xs.nblk * kT) {
    const uint32_t v = o2n[o];
    int64_t e = (int64_t)v < n_c ? run_entry(pred_new[v], epoch) : (int64_t)-1;
    if (e == -1 && has_parent && ((has_parent[o >> 6] >> (o & 63)) & 1ull)) {
      const uint32_t p = tree_parent[o];
      if (p != kNoParent) e = (int64_t)((uint64_t)p | (0xFFFFull << 48));
    }
    if (o == root) e = root;
    __builtin_nontemporal_store(e, pred + o);
  }
}

// Without CoreBFS, a vertex without edges is never reached and its pred
// entry is always -1: the engine fills pred once and the gather walks only
// core, the (original ID, new ID) pairs of the vertices with edges in
// ascending original-ID order.  One pair per thread: per vertex one
// random 8-byte read (its entry), the rest streams.  Started at the switch
// back to top-down (snap_level > 0: the level of that conversion), it writes
// only the vertices inside the bitmap snapshot -- this run's entries below
// that level, which nothing writes again -- and leaves the rest to
// k_pred_fixup, so the two write disjoint entries and need no order between
// them.  An entry the top-down levels write while the gather reads it is
// outside the snapshot whether it is read before (another epoch) or after
// (its level).
__global__ __launch_bounds__(kT) void k_pred_gather_core(const uint2* __restrict__ core,
                                                        int64_t n_core, uint32_t epoch,
                                                        const int64_t* __restrict__ pred_new,
                                                        int64_t root,
                                                        int64_t* __restrict__ pred,
                                                        uint32_t snap_level,
                                                        const LevelSync* ls, uint32_t snap_code) {
  if (!snapshot_is(ls, snap_code)) return;
  const XcdSpan xs = xcd_span(n_core);
  for (int64_t i = xs.lo + xs.rank * kT + threadIdx.x; i < xs.hi; i += xs.nblk * kT) {
    const uint2 ov = get_pair(core + i);
    const int64_t raw = pred_new[ov.y];
    if (snap_level != 0 && (run_entry(raw, epoch) == -1 ||
                            (uint32_t)((uint64_t)raw >> 48) >= snap_level))
      continue;
    int64_t e = run_entry(raw, epoch);
    if ((int64_t)ov.x == root) e = root;
    __builtin_nontemporal_store(e, pred + ov.x);
  }
}

__global__ __launch_bounds__(kT) void k_core_flags(const uint32_t* __restrict__ o2n, int64_t n,
                                                  int64_t n_c, uint32_t* __restrict__ flag) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (o < n) flag[o] = (int64_t)o2n[o] < n_c ? 1u : 0u;
}

__global__ __launch_bounds__(kT) void k_core_list(const uint32_t* __restrict__ flag,
                                                 const uint64_t* __restrict__ off,
                                                 const uint32_t* __restrict__ o2n, int64_t n,
                                                 uint2* __restrict__ core) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (o < n && flag[o]) core[off[o]] = make_uint2((uint32_t)o, o2n[o]);
}

// pred entries of the vertices outside the snapshot the early gather used:
// one thread per bitmap word, entries scattered by original ID.  With
// `outside` (the gather skipped those vertices) every vertex outside the
// snapshot gets its entry if the run reached it and -1 if not; otherwise
// (the gather wrote every vertex, and this runs after it) the ones reached
// after the snapshot (cur & ~snap) get their entries.  Queued on a guess
// that the run ends (done_code != 0), it returns at once unless the loop
// state says so.
__global__ __launch_bounds__(kT) void k_pred_fixup(const uint64_t* __restrict__ cur,
                                                  const uint64_t* __restrict__ snap,
                                                  int64_t words,
                                                  const uint32_t* __restrict__ n2o,
                                                  const int64_t* __restrict__ pred_new,
                                                  int64_t* __restrict__ pred, int outside,
                                                  const LevelSync* ls, uint32_t done_code) {
  if (done_code != 0 && !step_is_next(ls, done_code)) return;
  const int64_t w = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (w >= words) return;
  const uint64_t c = cur[w];
  uint64_t bits = outside ? ~snap[w] : c & ~snap[w];
  while (bits) {
    const int b = __ffsll((long long)bits) - 1;
    bits &= bits - 1;
    const uint32_t v = (uint32_t)(w * 64 + b);
    pred[n2o[v]] = ((c >> b) & 1ull) ? (int64_t)((uint64_t)pred_new[v] & ~kEpochMask)
                                     : (int64_t)-1;
  }
}

// The walked path from the root to the core, with real depths.
constexpr int kMaxPathArgs = 32;
struct PathArgs {
  int count;
  uint32_t v[kMaxPathArgs];
  int64_t e[kMaxPathArgs];
};

__global__ void k_write_path(PathArgs pa, int64_t* __restrict__ pred) {
  const int i = threadIdx.x;
  if (i < pa.count) pred[pa.v[i]] = pa.e[i];
}

__global__ __launch_bounds__(kT) void k_levels(const int64_t* __restrict__ pred,
                                              int64_t n,
                                              uint16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < n) out[i] = (uint16_t)((uint64_t)pred[i] >> 48);
}

unsigned grid_for(int64_t work_items, int64_t per_block) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(kMaxGrid, ceil_div(work_items, per_block)));
}

}  // namespace

void bfs_init(g500_bfs* b, int edgefactor, double alpha, double beta) {
  g500_graph* g = b->g;
  b->words = ceil_div(std::max<int64_t>(g->n_c, 1), 64);
  b->pred.alloc((size_t)g->n);
  b->pred_new.alloc((size_t)std::max<int64_t>(g->n_c, 1));
  // every entry starts with the epoch no run uses (run_entry)
  G500_CHECK_HIP(hipMemset(b->pred_new.get(), 0xFF, b->pred_new.n * sizeof(int64_t)));
  b->epoch = 0xFFFFFFFFu;  // the first run's epoch is 0
  b->vis.alloc((size_t)b->words);
  b->vis_alt.alloc((size_t)b->words);
  b->q.alloc((size_t)std::max<int64_t>(g->n_c, 1));
  b->nq.alloc((size_t)std::max<int64_t>(g->n_c, 1));
  b->dsum.alloc((size_t)std::max<int64_t>(g->n_c, 1) + 1);
  // a frontier's edges never exceed m
  b->tile_owner.alloc((size_t)ceil_div(std::max<int64_t>(g->m, 1), kTdTile) + 2);
  b->bsum.alloc((size_t)ceil_div(std::max<int64_t>(g->n_c, 1), kPrepTile));
  b->boff.alloc((size_t)ceil_div(std::max<int64_t>(g->n_c, 1), kPrepTile));
  b->prep_ticket.alloc(1);
  G500_CHECK_HIP(hipMemset(b->prep_ticket.get(), 0, sizeof(unsigned int)));
  const int64_t chunks = ceil_div(b->words, kWave);
  b->unv.alloc((size_t)(chunks * kChunkVerts));
  b->unv_cnt.alloc((size_t)chunks);
  // bottom-up grids: what stays resident (each block stages the hub bits
  // once), never more blocks than chunk groups
  int cus = 0, per_cu[3] = {0, 0, 0};
  G500_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
  G500_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu[0], reinterpret_cast<const void*>(&k_bu_sweep), kBuT, 0));
  G500_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu[1], reinterpret_cast<const void*>(&k_bu_list), kBuT, 0));
  G500_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu[2], reinterpret_cast<const void*>(&k_bu_long), kBuT, 0));
  for (int i = 0; i < 3; ++i)
    b->bu_grid[i] = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)std::max(cus, 1) * std::max(per_cu[i], 1),
                             ceil_div(chunks, kBuWaves)));
  b->lpart.alloc((size_t)std::max<int64_t>(
      kMaxGrid, std::max<int64_t>(b->bu_grid[0] + b->bu_grid[2], b->bu_grid[1])));
  b->wacc.alloc((size_t)b->bu_grid[0] + b->bu_grid[2] + b->bu_grid[1]);
  b->lsync.alloc(1);
  G500_CHECK_HIP(hipMemset(b->lsync.get(), 0, sizeof(LevelSync)));
  G500_CHECK_HIP(hipHostMalloc((void**)&b->h_report, kReportSlots * sizeof(HostReport),
                               hipHostMallocMapped | hipHostMallocCoherent));
  std::memset((void*)b->h_report, 0, kReportSlots * sizeof(HostReport));
  G500_CHECK_HIP(hipHostGetDevicePointer((void**)&b->d_report, b->h_report, 0));
  // ev_begin only times the run; ev_end also tells the host the run is done
  G500_CHECK_HIP(hipEventCreateWithFlags(&b->ev_begin, hipEventDisableSystemFence));
  G500_CHECK_HIP(hipEventCreate(&b->ev_end));
  G500_CHECK_HIP(hipEventCreateWithFlags(&b->ev_snap, hipEventDisableTiming));
  G500_CHECK_HIP(hipEventCreateWithFlags(&b->ev_g0, hipEventDisableSystemFence));
  G500_CHECK_HIP(hipEventCreate(&b->ev_g1));
  G500_CHECK_HIP(hipStreamCreate(&b->gstream));
  b->vis_snap.alloc((size_t)b->words);
  b->qcnt.alloc((size_t)kQueueGrid * kWavesPerBlock);
  b->qedges.alloc((size_t)kQueueGrid);
  if (!g->corebfs && g->n_c > 0) {
    // the core's original IDs in ascending order; pred starts at -1, which
    // the vertices without edges keep
    hipStream_t s = b->stream;
    DevBuf<uint32_t> flag;
    DevBuf<uint64_t> off;
    flag.alloc((size_t)g->n);
    off.alloc((size_t)g->n + 1);
    hipLaunchKernelGGL(k_core_flags, dim3((unsigned)ceil_div(g->n, kT)), dim3(kT), 0, s,
                       g->orig2new.get(), g->n, g->n_c, flag.get());
    G500_CHECK_LAUNCH();
    ScanWorkspace ws;
    exclusive_scan<uint32_t>(flag.get(), off.get(), g->n, ws, s);
    b->core.alloc((size_t)g->n_c);
    hipLaunchKernelGGL(k_core_list, dim3((unsigned)ceil_div(g->n, kT)), dim3(kT), 0, s,
                       flag.get(), off.get(), g->orig2new.get(), g->n, b->core.get());
    G500_CHECK_LAUNCH();
    uint64_t total = 0;
    G500_CHECK_HIP(hipMemcpyAsync(&total, off.get() + g->n, sizeof(total),
                                  hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
    if ((int64_t)total != g->n_c)
      fail(G500_ERR_STATE, "core list holds %llu vertices, expected %lld",
           (unsigned long long)total, (long long)g->n_c);
    b->n_core = g->n_c;
    b->pred_fill_needed = true;
  }
  ensure_host_orig2new(g);
  if (const char* ea = getenv("G500_EXEC_ALPHA")) b->exec_alpha = atof(ea);
  if (const char* sp = getenv("G500_SPECULATE")) b->spec_depth = std::max(0, atoi(sp));
  if (const char* gw = getenv("G500_GUESS")) b->guess_wrong = std::strcmp(gw, "wrong") == 0;
  if (g->corebfs && !g->core_reset_done) {
    // corebfs_index::reset_unreachable: one BFS over the core from its
    // highest-degree vertex decides which forest parents survive
    if (g->hd_root >= 0) {
      bfs_run(b, g->hd_root, edgefactor, alpha, beta, nullptr);
      corebfs_reset_unreachable(g, b->vis.get(), b->stream);
    } else {
      corebfs_reset_unreachable(g, nullptr, b->stream);
    }
    b->have_result = false;
  }
}

namespace {

// Spins until the GPU posted report `seq` (slot seq % kReportSlots); returns
// its four values.
void wait_report(g500_bfs* b, unsigned long long seq, unsigned long long out[4]) {
  HostReport* hr = b->h_report + (seq % kReportSlots);
  const unsigned long long tag = seq & ((1ull << kReportTagBits) - 1);
  const unsigned long long tmask = (1ull << kReportTagBits) - 1;
  auto whole = [&]() {
    for (int i = 0; i < 4; ++i) {
      const unsigned long long w = __atomic_load_n(&hr->w[i], __ATOMIC_ACQUIRE);
      if ((w & tmask) != tag) return false;
      out[i] = w >> kReportTagBits;
    }
    return true;
  };
  // The stream is queried only once a report is overdue: a query of a
  // stream with kernels in flight queues a cache-flush marker behind them,
  // which costs ~5.6 us of device time at the next step boundary.
  uint64_t spins = 0;
  const auto t0 = std::chrono::steady_clock::now();
  auto next_query = t0 + std::chrono::milliseconds(20);
  while (!whole()) {
    if ((++spins & 1023) == 0) {
      const auto now = std::chrono::steady_clock::now();
      if (now < next_query) continue;
      next_query = now + std::chrono::milliseconds(20);
      const hipError_t q = hipStreamQuery(b->stream);
      if (q == hipSuccess) {
        if (whole()) break;
        fail(G500_ERR_STATE, "step report %llu missing after the stream drained", seq);
      }
      if (q != hipErrorNotReady)
        fail(G500_ERR_HIP, "stream error while waiting for a level: %s",
             hipGetErrorString(q));
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300))
        fail(G500_ERR_STATE, "timed out waiting for step report %llu", seq);
    }
  }
}

// pred by original ID from this run's level entries (a smaller grid beside
// the last top-down levels measured slower: the random reads need every
// wave the chip holds).  snap_code != 0: queued before the conversion of
// that code is known to run (the kernel checks).
void launch_gather(g500_bfs* b, int64_t root, hipStream_t st, uint32_t snap_code,
                   uint32_t snap_level) {
  g500_graph* g = b->g;
  G500_CHECK_HIP(hipEventRecord(b->ev_g0, st));
  // grids are a multiple of kXcds (xcd_span)
  auto grid = [](int64_t items) {
    return (unsigned)std::min<int64_t>(ceil_div(ceil_div(items, kT), kXcds) * kXcds, 65536);
  };
  if (b->n_core > 0)
    hipLaunchKernelGGL(k_pred_gather_core, dim3(grid(b->n_core)), dim3(kT), 0, st,
                       b->core.get(), b->n_core, b->epoch, b->pred_new.get(), root,
                       b->pred.get(), snap_level, b->lsync.get(), snap_code);
  else
    hipLaunchKernelGGL(k_pred_gather, dim3(grid(g->n)),
                       dim3(kT), 0, st, g->orig2new.get(), b->epoch, b->pred_new.get(), g->n,
                       g->n_c, root, g->corebfs ? g->tree_parent.get() : nullptr,
                       g->corebfs ? g->has_parent.get() : nullptr, b->pred.get(),
                       b->lsync.get(), snap_code);
  G500_CHECK_LAUNCH();
  G500_CHECK_HIP(hipEventRecord(b->ev_g1, st));
  if (st != b->stream) b->gather_pending = true;
}

// The buffers a step starts from: the visited bitmap, the other bitmap (the
// one before the last level), the frontier queue and the next queue.
struct Bufs {
  uint64_t* vis;
  uint64_t* vis_alt;
  uint32_t* q;
  uint32_t* nq;
};

// A queued step: its code, report number, the buffers after it (if it
// runs) and its event pairs.
struct Step {
  uint32_t code = 0;
  unsigned long long seq = 0;
  Bufs after{};
  int ev = -1;        // level_events pair (timing runs)
  int sweep_ev = -1;  // sweep_events pair (steps with a sweep)
};

// Timing events: read only by hipEventElapsedTime after the run's end event
// (a default event) completed, so they skip the system-scope fence, which
// costs ~1.4 us of the ~6 us each record adds to its stream.
hipEvent_t pool_event(std::vector<hipEvent_t>& pool, size_t i) {
  while (pool.size() <= i) {
    hipEvent_t e;
    G500_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    pool.push_back(e);
  }
  return pool[i];
}

// Queues step `code` behind whatever the engine stream holds; `in` are the
// buffers it starts from.  Every kernel of the step returns at once unless
// the level loop's state names this step next.  A bottom-up level writes
// the other bitmap (the two swap roles) and lists its long rows in nq, which
// no bottom-up level uses (a top-down level run as a sweep refills it from
// the bitmap afterwards); a top-down level swaps the queues; the conversion
// at the switch to top-down writes the queue and the gather's bitmap
// snapshot (ev_snap is recorded as soon as that is written).
Step launch_step(g500_bfs* b, uint32_t code, const Bufs& in, size_t* nlev_ev,
                 size_t* nsweep_ev) {
  g500_graph* g = b->g;
  hipStream_t s = b->stream;
  LevelSync* ls = b->lsync.get();
  Step st;
  st.code = code;
  st.seq = ++b->seq;
  st.after = in;
  const uint32_t mode = step_mode(code);
  const uint64_t level_bits = ((uint64_t)step_level(code) << 48) | ((uint64_t)b->epoch << 32);
  if (b->timing) {
    st.ev = (int)*nlev_ev;
    *nlev_ev += 2;
    G500_CHECK_HIP(hipEventRecord(pool_event(b->level_events, st.ev), s));
  }
  // counting runs: per-block scan bytes, sweep / long-row / list slots
  uint64_t* wsweep = b->counting ? b->wacc.get() : nullptr;
  uint64_t* wlong = b->counting ? b->wacc.get() + b->bu_grid[0] : nullptr;
  uint64_t* wlist = b->counting ? b->wacc.get() + b->bu_grid[0] + b->bu_grid[2] : nullptr;
  auto sweep = [&]() {
    st.sweep_ev = (int)*nsweep_ev;
    *nsweep_ev += 2;
    G500_CHECK_HIP(hipEventRecord(pool_event(b->sweep_events, st.sweep_ev), s));
    hipLaunchKernelGGL(k_bu_sweep, dim3(b->bu_grid[0]), dim3(kBuT), 0, s, in.vis, in.vis_alt,
                       b->words, g->row_off.get(), g->col.get(), g->first.get(),
                       g->multi.get(), g->new2orig.get(), b->pred_new.get(), b->unv.get(),
                       b->unv_cnt.get(), in.nq, ls, code, b->lpart.get(), level_bits, wsweep);
    G500_CHECK_LAUNCH();
    G500_CHECK_HIP(hipEventRecord(pool_event(b->sweep_events, st.sweep_ev + 1), s));
    hipLaunchKernelGGL(k_bu_long, dim3(b->bu_grid[2]), dim3(kBuT), 0, s, in.vis,
                       (unsigned long long*)in.vis_alt, b->words, g->row_off.get(),
                       g->col.get(), g->new2orig.get(), b->pred_new.get(), b->unv.get(),
                       b->unv_cnt.get(), in.nq, ls, code, b->lpart.get() + b->bu_grid[0],
                       level_bits, wlong);
    G500_CHECK_LAUNCH();
    st.after.vis = in.vis_alt;
    st.after.vis_alt = in.vis;
  };
  // bitmap -> queue: the bits set in cur but not in prev
  auto to_queue = [&](const uint64_t* cur, const uint64_t* prev, uint32_t* out,
                      uint64_t* snap) {
    const dim3 grid((unsigned)std::min<int64_t>(kQueueGrid, ceil_div(b->words, (int64_t)kT)));
    hipLaunchKernelGGL(k_queue_count, grid, dim3(kT), 0, s, cur, prev, b->words, b->qcnt.get(),
                       snap, ls, code);
    G500_CHECK_LAUNCH();
    if (snap) G500_CHECK_HIP(hipEventRecord(b->ev_snap, s));
    hipLaunchKernelGGL(k_queue_write, grid, dim3(kT), 0, s, cur, prev, b->words,
                       g->deg_new.get(), b->qcnt.get(), out, b->qedges.get(), ls, code);
    G500_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_queue_decide, dim3(1), dim3(kT), 0, s, b->qcnt.get(), b->qedges.get(),
                       (int)grid.x, ls, b->d_report, st.seq, code);
    G500_CHECK_LAUNCH();
  };
  switch (mode) {
    case kModeTD:
    case kModeTDS:
      if (mode == kModeTD) {
        hipLaunchKernelGGL(k_td_prep_sums, dim3(kPrepGrid), dim3(kT), 0, s, in.q, ls, code,
                           g->row_off.get(), b->dsum.get(), b->bsum.get(), b->boff.get(),
                           b->prep_ticket.get());
        G500_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_td_prep_tiles, dim3(kPrepGrid), dim3(kT), 0, s, ls, code,
                           b->dsum.get(), b->bsum.get(), b->boff.get(), b->tile_owner.get());
        G500_CHECK_LAUNCH();
      }
      hipLaunchKernelGGL(k_td_expand, dim3(kMaxGrid), dim3(kT), 0, s, in.q, code,
                         b->dsum.get(), b->tile_owner.get(), g->row_off.get(), g->col.get(),
                         g->new2orig.get(), g->deg_new.get(), (unsigned long long*)in.vis,
                         b->pred_new.get(), in.nq, ls, b->lpart.get(), level_bits);
      G500_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_level_decide, dim3(1), dim3(kT), 0, s, (const uint64_t*)nullptr, 0,
                         b->lpart.get(), kMaxGrid, ls, b->d_report, st.seq, code);
      G500_CHECK_LAUNCH();
      std::swap(st.after.q, st.after.nq);
      break;
    case kModeTDX:
      sweep();
      to_queue(in.vis_alt, in.vis, in.nq, nullptr);
      std::swap(st.after.q, st.after.nq);
      break;
    case kModeBUSweep:
      sweep();
      hipLaunchKernelGGL(k_level_decide, dim3(1), dim3(kT), 0, s, b->lpart.get(),
                         (int)(b->bu_grid[0] + b->bu_grid[2]), (const uint64_t*)nullptr, 0, ls,
                         b->d_report, st.seq, code);
      G500_CHECK_LAUNCH();
      break;
    case kModeBUList:
      hipLaunchKernelGGL(k_bu_list, dim3(b->bu_grid[1]), dim3(kBuT), 0, s, in.vis, in.vis_alt,
                         b->words, g->row_off.get(), g->col.get(), g->first.get(),
                         g->multi.get(), g->new2orig.get(), b->pred_new.get(), b->unv.get(),
                         b->unv_cnt.get(), ls, code, b->lpart.get(), level_bits, wlist);
      G500_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_level_decide, dim3(1), dim3(kT), 0, s, b->lpart.get(),
                         (int)b->bu_grid[1], (const uint64_t*)nullptr, 0, ls, b->d_report,
                         st.seq, code);
      G500_CHECK_LAUNCH();
      st.after.vis = in.vis_alt;
      st.after.vis_alt = in.vis;
      break;
    case kModeConv:
      to_queue(in.vis, in.vis_alt, in.q, b->vis_snap.get());
      break;
    default:
      fail(G500_ERR_STATE, "no step of mode %u", mode);
  }
  if (b->timing) G500_CHECK_HIP(hipEventRecord(pool_event(b->level_events, st.ev + 1), s));
  return st;
}

// Host mirror of the level loop (mpi/bfs.hpp:3264-3279, decision at
// 3413-3457): the same rule as decide_next on the device, applied to each
// step's reported totals; it fills auto_tuning_data and the trace and
// checks the device's choice.
struct HostLoop {
  bool growing = true;
  int64_t prev_nq = 0;
  int64_t alpha_thr = 0, beta_thr = 0, exec_thr = 0;
  int64_t n_c = 0, m = 0;
  // for the guesses: vertices with edges reached so far, and how many of
  // them the engine's previous run left unreached (other components)
  int64_t visited = 0, unreach = 0;

  // next step after level step `code` reported (nq_size, nq_edges)
  uint32_t after_level(uint32_t code, int64_t nq_size, int64_t nq_edges, int64_t* at) {
    const uint32_t mode = step_mode(code);
    const int level = step_level(code);
    const bool td = mode_is_td(mode);
    visited += nq_size;
    if (nq_size == 0) return step_code(level + 1, kModeDone);
    bool next_td = td;
    if (growing && nq_size > prev_nq) {
      if (td) {
        if (nq_edges > alpha_thr) {
          next_td = false;
          at[G500_AT_GLOBAL_NQ_EDGES] = nq_edges;
          at[G500_AT_NUM_GLOBAL_EDGES] = m;
          at[G500_AT_T2B_LEVEL] = level;
        } else {
          at[G500_AT_PRE_GLOBAL_NQ_EDGES] = nq_edges;
        }
      }
    } else if (!td) {
      if (nq_size < beta_thr) {
        next_td = true;
        growing = false;
        at[G500_AT_GLOBAL_NQ_SIZE] = nq_size;
        at[G500_AT_NUM_GLOBAL_VERTS] = n_c;
        at[G500_AT_B2T_LEVEL] = level;
      } else {
        at[G500_AT_PRE_GLOBAL_NQ_SIZE] = nq_size;
      }
    }
    prev_nq = nq_size;
    uint32_t next;
    if (td && next_td) next = td_step_mode(nq_size, nq_edges, exec_thr);
    else if (td) next = mode == kModeTDX ? kModeBUList : kModeBUSweep;
    else next = next_td ? kModeConv : kModeBUList;
    return step_code(level + 1, next);
  }
};

// Frontier totals (size, degree sum) that a level of mode `mode` is expected
// to find, from the last level of `tr` (reported or itself estimated).  A
// top-down level over a frontier of degree sum fe0 finds about fe0 new
// vertices (early on, most edges lead to unvisited ones; none once every
// vertex outside the other components is reached), whose degree sum
// on Kronecker graphs is about 24 sqrt(m fe0) (the frontier reaches the hubs
// at once: fitted to SCALE-24 and 26 runs, within a factor of two from fe0
// ~ 10^3 to 10^7).  A bottom-up level finds the last level's count times its
// growth, but at most 0.8 of the vertices still unreached that are not in
// another component (learned from the previous run).
void estimate_level(uint32_t mode, const std::vector<g500_level_trace>& tr, const HostLoop& hl,
                    int64_t* nq, int64_t* fe) {
  const size_t k = tr.size();
  const double left = (double)std::max<int64_t>(0, hl.n_c - hl.visited - hl.unreach);
  if (mode_is_td(mode)) {
    const double fe0 = k >= 1 ? (double)tr[k - 1].nq_edges : 0.0;
    *nq = (int64_t)std::min(std::max(fe0, 1.0), left);
    *fe = (int64_t)std::min((double)hl.m, 24.0 * std::sqrt((double)hl.m * std::max(fe0, 1.0)));
    return;
  }
  const double nq0 = k >= 1 ? (double)tr[k - 1].nq_size : 1.0;
  const double g = (k >= 2 && tr[k - 2].nq_size > 0) ? nq0 / (double)tr[k - 2].nq_size : 1.0;
  *nq = (int64_t)std::max(1.0, std::min(nq0 * g, 0.8 * left));
  *fe = 0;
}

// The step the host queues behind `last` before `last` has reported: the
// direction rule (a copy of the host mirror) applied to the totals `last` is
// expected to find.  `tr` holds the reported levels plus the expected ones
// of the steps queued before `last`.  A wrong guess costs the guessed step's
// empty launches and one host round trip, never a result.
uint32_t predict_next(const Step& last, HostLoop hl, const std::vector<g500_level_trace>& tr) {
  const uint32_t mode = step_mode(last.code);
  // the conversion's frontier is the last level's
  if (mode == kModeConv)
    return step_code(step_level(last.code),
                     td_step_mode(tr.empty() ? 0 : tr.back().nq_size, 0, hl.exec_thr));
  int64_t nq, fe, at[G500_AT_NUM];
  estimate_level(mode, tr, hl, &nq, &fe);
  return hl.after_level(last.code, nq, fe, at);
}

// The end of a run queued on guesses: the pred gather behind a guessed
// conversion (k_pred_gather* checks that the conversion ran) and the fixup
// behind the step guessed to be the last (k_pred_fixup checks that the loop
// ended there).
struct EndGuess {
  int64_t root = 0;
  bool gather_confirmed = false;     // a gather follows the conversion that ran
  unsigned long long conv_seq = 0;   // the guessed conversion the last guessed gather follows
  unsigned long long fixup_seq = 0;  // the step the guessed fixup follows
  uint32_t fixup_code = 0;           // the final code it waits for
};

// The engine stream waits until every gather queued so far has ended.
void wait_gathers(g500_bfs* b) {
  G500_CHECK_HIP(hipStreamWaitEvent(b->stream, b->ev_g1, 0));
  b->gather_pending = false;
}

// Without CoreBFS the early gather writes only the vertices inside the
// bitmap snapshot and the fixup all the others, so the fixup need not wait
// for the gather (a handoff between streams costs ~25 us); the host waits
// for both streams at the end of the run.  With CoreBFS the gather writes
// every entry (forest parents included) and the fixup follows it.
bool disjoint_fixup(const g500_bfs* b) { return b->n_core > 0; }

// The gather on gstream behind the snapshot of the conversion queued last
// (ev_snap; conv_code: that conversion's code; snap_code: the same when it
// is queued on a guess, else 0); ev_g1, which launch_gather records behind
// it, is what the engine stream waits on.
void queue_gather(g500_bfs* b, int64_t root, uint32_t snap_code, uint32_t conv_code) {
  G500_CHECK_HIP(hipStreamWaitEvent(b->gstream, b->ev_snap, 0));
  launch_gather(b, root, b->gstream, snap_code,
                disjoint_fixup(b) ? (uint32_t)step_level(conv_code) : 0u);
}

void queue_fixup(g500_bfs* b, const uint64_t* cur, uint32_t done_code) {
  hipStream_t s = b->stream;
  const bool outside = disjoint_fixup(b);
  if (!outside) wait_gathers(b);
  hipLaunchKernelGGL(k_pred_fixup, dim3((unsigned)ceil_div(b->words, kT)), dim3(kT), 0, s, cur,
                     b->vis_snap.get(), b->words, b->g->new2orig.get(), b->pred_new.get(),
                     b->pred.get(), outside ? 1 : 0, b->lsync.get(), done_code);
  G500_CHECK_LAUNCH();
}

// The test hook's wrong guess for a guessed step `code`: a guessed end of the
// run becomes a top-down level, a conversion a list level, a bottom-up level
// a conversion (its gather is queued too) and a top-down level the end of
// the run (the fixup is queued too, once the gather runs).
uint32_t wrong_guess(uint32_t code) {
  const int l = step_level(code);
  switch (step_mode(code)) {
    case kModeDone: return step_code(l, kModeTD);
    case kModeConv: return step_code(l, kModeBUList);
    case kModeBUList:
    case kModeBUSweep: return step_code(l, kModeConv);
    default: return step_code(l, kModeDone);
  }
}

// Queues guessed steps behind the queued ones until `depth` are queued:
// the steps queued after the first one (whose predecessor has reported)
// contribute their expected totals to the guess of the next one.  A guessed
// conversion gets the pred gather behind it on the second stream; a step
// guessed to end the run gets the fixup behind it once the gather is
// confirmed.
void speculate(g500_bfs* b, std::deque<Step>& pend, const HostLoop& hl, size_t depth,
               size_t* nlev_ev, size_t* nsweep_ev, EndGuess* eg) {
  if (pend.empty() || pend.size() >= depth) return;
  HostLoop h = hl;
  std::vector<g500_level_trace> tr = b->trace;
  for (size_t i = 0; i + 1 < pend.size(); ++i) {
    const uint32_t mode = step_mode(pend[i].code);
    if (mode == kModeConv) continue;
    g500_level_trace t{};
    int64_t at[G500_AT_NUM];
    estimate_level(mode, tr, h, &t.nq_size, &t.nq_edges);
    t.level = step_level(pend[i].code);
    t.top_down = mode_is_td(mode) ? 1 : 0;
    (void)h.after_level(pend[i].code, t.nq_size, t.nq_edges, at);
    tr.push_back(t);
  }
  while (pend.size() < depth) {
    const Step& last = pend.back();
    uint32_t next = predict_next(last, h, tr);
    if (b->guess_wrong) next = wrong_guess(next);
    if (step_mode(next) == kModeDone) {
      if (eg->gather_confirmed && eg->fixup_seq != last.seq) {
        queue_fixup(b, last.after.vis, next);
        eg->fixup_seq = last.seq;
        eg->fixup_code = next;
      }
      return;
    }
    const uint32_t lm = step_mode(last.code);
    if (lm != kModeConv) {
      g500_level_trace t{};
      int64_t at[G500_AT_NUM];
      estimate_level(lm, tr, h, &t.nq_size, &t.nq_edges);
      t.level = step_level(last.code);
      t.top_down = mode_is_td(lm) ? 1 : 0;
      (void)h.after_level(last.code, t.nq_size, t.nq_edges, at);
      tr.push_back(t);
    }
    pend.push_back(launch_step(b, next, last.after, nlev_ev, nsweep_ev));
    if (step_mode(next) == kModeConv && !eg->gather_confirmed) {
      // ev_snap was just recorded behind this conversion's snapshot
      queue_gather(b, eg->root, next, next);
      eg->conv_seq = pend.back().seq;
    }
  }
}

}  // namespace

void bfs_run(g500_bfs* b, int64_t root, int edgefactor, double alpha,
             double beta, int64_t* at) {
  g500_graph* g = b->g;
  hipStream_t s = b->stream;
  if (root < 0 || root >= g->n)
    fail(G500_ERR_ARG, "root %lld outside [0, %lld)", (long long)root, (long long)g->n);
  if (!(alpha > 0) || !(beta > 0) || edgefactor <= 0)
    fail(G500_ERR_ARG, "alpha, beta and edgefactor must be positive");
  int64_t at_local[G500_AT_NUM];
  if (!at) {
    for (int i = 0; i < G500_AT_NUM; ++i) at_local[i] = -1;
    at = at_local;
  }
  const int64_t n_c = g->n_c, words = b->words;
  // CoreBFS: walk up the forest to the core (bfs_tree_with_callback,
  // external/indexed_bfs/bfs/corebfs.hpp:125-157); the path vertices get
  // parents pointing back toward the root and their true depths
  std::vector<std::pair<uint32_t, int64_t>> path;
  int64_t bfs_root = root;
  int root_level = 0;
  if (g->corebfs) {
    int64_t u = root, p = root;
    for (int lv = 0;; ++lv) {
      if (lv > (1 << 20)) fail(G500_ERR_STATE, "forest walk from %lld does not end", (long long)root);
      path.emplace_back((uint32_t)p, (int64_t)(((uint64_t)lv << 48) | (uint64_t)u));
      u = p;
      p = (has_parent_host(g, u) && !g->h_tree_parent.empty() &&
           g->h_tree_parent[(size_t)u] != kNoParent)
              ? (int64_t)g->h_tree_parent[(size_t)u]
              : -1;
      if (p < 0) {
        bfs_root = u;
        root_level = lv;
        break;
      }
    }
  }
  const int64_t root_new = g->h_orig2new.empty() ? -1 : (int64_t)g->h_orig2new[bfs_root];
  const bool root_in = root_new >= 0 && root_new < n_c;

  b->trace.clear();
  b->stats = g500_run_stats{};
  b->have_result = false;
  // this run's epoch; before the epochs wrap, every entry is set to the
  // epoch no run uses, so no entry of an earlier run can match
  if (++b->epoch >= kEpochNone) {
    if (b->gather_pending) wait_gathers(b);
    G500_CHECK_HIP(hipMemsetAsync(b->pred_new.get(), 0xFF, b->pred_new.n * sizeof(int64_t), s));
    b->epoch = 0;
  }
  // a gather left running by a run that failed must not overlap this one
  // (after a run that ended normally the host or the fixup waited for it)
  if (b->gather_pending) wait_gathers(b);
  if (b->n_core > 0 && b->pred_fill_needed) {
    // entries of vertices without edges are -1 and stay so; the fill is
    // redone only after the buffer was overwritten (validation of a host pred)
    G500_CHECK_HIP(hipMemsetAsync(b->pred.get(), 0xFF, (size_t)g->n * sizeof(int64_t), s));
    b->pred_fill_needed = false;
    b->iso_root = -1;
  }
  // counting runs: the bytes each kernel needs (closed form per level from
  // the level totals, plus the row-scan bytes the bottom-up kernels count)
  int64_t alg = 0, sweep_host = 0, root_deg = 0;
  if (b->counting) {
    G500_CHECK_HIP(hipMemsetAsync(b->wacc.get(), 0, b->wacc.n * sizeof(uint64_t), s));
    if (root_in) {
      int64_t ro[2];
      G500_CHECK_HIP(hipMemcpyAsync(ro, g->row_off.get() + root_new, sizeof(ro),
                                    hipMemcpyDeviceToHost, s));
      G500_CHECK_HIP(hipStreamSynchronize(s));
      root_deg = ro[1] - ro[0];
    }
  }
  int64_t visited = root_in ? 1 : 0;  // vertices with edges reached so far
  int64_t after_snap = 0;             // of them, reached after the gather's snapshot
  alg += 8 * words;                   // init writes the bitmap

  HostLoop hl;
  hl.n_c = n_c;
  hl.m = g->m;
  hl.alpha_thr = (int64_t)((double)g->m / alpha);
  hl.beta_thr = (int64_t)((double)n_c / (beta * edgefactor * 2.0));
  hl.exec_thr = b->exec_alpha > 0 ? (int64_t)std::floor((double)g->m / b->exec_alpha)
                                  : INT64_MAX;
  hl.visited = root_in ? 1 : 0;
  hl.unreach = b->unreach;
  const uint32_t first = step_code(root_level + 1, kModeTDS);
  LoopInit li{first, root_in ? 1 : 0, 0, hl.alpha_thr, hl.beta_thr, hl.exec_thr};
  G500_CHECK_HIP(hipEventRecord(b->ev_begin, s));
  hipLaunchKernelGGL(k_bfs_init, dim3(grid_for(words, kT)), dim3(kT), 0, s,
                     b->vis.get(), words, n_c, root_in ? root_new : (int64_t)-1,
                     b->pred_new.get(),
                     (int64_t)(((uint64_t)root_level << 48) | ((uint64_t)b->epoch << 32) |
                               (uint64_t)bfs_root),
                     b->q.get(),
                     b->lsync.get(), li, g->row_off.get());
  G500_CHECK_LAUNCH();

  int64_t qlen = root_in ? 1 : 0;
  int64_t est_edges = -1;  // degree sum of the frontier (unknown for the root)
  bool early_gather = false;  // pred gather already running on gstream
  EndGuess eg;
  eg.root = root;
  uint32_t final_code = 0;
  int64_t last_nq = 0;
  unsigned long long rep[4];
  size_t nlev_ev = 0, nsweep_ev = 0;
  std::deque<Step> pend;
  std::vector<Step> ran;  // steps that ran, in order
  int mispredicted = 0;
  pend.push_back(launch_step(b, first, Bufs{b->vis.get(), b->vis_alt.get(), b->q.get(),
                                            b->nq.get()},
                             &nlev_ev, &nsweep_ev));
  Bufs cur{};
  for (;;) {
    if (b->spec_depth > 0)
      speculate(b, pend, hl, (size_t)b->spec_depth + 1, &nlev_ev, &nsweep_ev, &eg);
    const Step st = pend.front();
    pend.pop_front();
    wait_report(b, st.seq, rep);
    ran.push_back(st);
    cur = st.after;
    const uint32_t mode = step_mode(st.code);
    uint32_t next;
    if (mode == kModeConv) {
      // the frontier of the first top-down level after the switch
      qlen = (int64_t)rep[0];
      est_edges = (int64_t)rep[1];
      if (qlen != last_nq)
        fail(G500_ERR_STATE, "frontier conversion mismatch (%lld vs %lld)", (long long)qlen,
             (long long)last_nq);
      alg += 24 * words + 8 * qlen;  // bitmap -> queue with the snapshot
      next = step_code(step_level(st.code), td_step_mode(qlen, est_edges, hl.exec_thr));
    } else {
      const int level = step_level(st.code);
      const bool td = mode_is_td(mode);
      at[G500_AT_LEVEL] = level;
      const int64_t nq_size = (int64_t)rep[0];
      const int64_t nq_edges = td ? (int64_t)rep[1] : 0;
      g500_level_trace tr{};
      tr.level = level;
      tr.top_down = td ? 1 : 0;
      tr.nq_size = nq_size;
      tr.nq_edges = nq_edges;
      tr.exec_bottom_up = (mode == kModeTD || mode == kModeTDS) ? 0 : 1;
      b->trace.push_back(tr);
      {
        // bytes this level needs beyond the bottom-up kernels' counted scans
        const int64_t unvisited = n_c - visited;
        if (mode == kModeTDX || mode == kModeBUSweep) {
          // sweep: old bitmap, multi bitmap and new bitmap; a first-neighbour
          // pair per unvisited vertex
          const int64_t sw = 24 * words + 8 * unvisited;
          sweep_host += sw;
          alg += sw;
          if (td) alg += 16 * words + 8 * nq_size;  // bitmap -> queue (queue + degree)
        } else if (mode == kModeBUList) {
          // list level: old and new bitmap; a list entry and a pair per listed vertex
          alg += 16 * words + 12 * unvisited;
        } else {
          // top-down: queue entry, row offsets and original ID per frontier
          // vertex, one column entry per frontier edge, and per vertex found its
          // entry, queue slot and degree (plus the prefix sums of a large frontier)
          const int64_t fe = est_edges < 0 ? root_deg : est_edges;
          alg += 24 * qlen + 4 * fe + 16 * nq_size + (qlen > kTdWin ? 16 * qlen : 0);
        }
        visited += nq_size;
        if (early_gather) after_snap += nq_size;
      }
      next = hl.after_level(st.code, nq_size, nq_edges, at);
      last_nq = nq_size;
      if (td && nq_size > 0) {
        qlen = nq_size;
        est_edges = nq_edges;
      }
    }
    if ((uint32_t)rep[2] != next)
      fail(G500_ERR_STATE, "level %d: the device chose step %u:%u, the reference rule %u:%u",
           step_level(st.code), step_level((uint32_t)rep[2]), step_mode((uint32_t)rep[2]),
           step_level(next), step_mode(next));
    if (step_mode(next) == kModeDone) {
      final_code = next;
      break;
    }
    if (pend.empty() || pend.front().code != next) {
      // predicted wrong: the queued steps return at once
      mispredicted += (int)pend.size();
      pend.clear();
      pend.push_back(launch_step(b, next, cur, &nlev_ev, &nsweep_ev));
    }
    if (step_mode(next) == kModeConv) {
      // switching to top-down: the remaining top-down levels are short and
      // latency-bound, so the pred gather of every vertex reached so far
      // starts on the second stream once the conversion has written the
      // bitmap snapshot (ev_snap, recorded when the conversion was queued);
      // the vertices the last levels reach are patched in at the end
      // (k_pred_fixup).  A conversion queued on a guess has its gather
      // queued behind it already.
      if (eg.conv_seq != pend.front().seq) queue_gather(b, root, 0, next);
      early_gather = true;
      eg.gather_confirmed = true;
    }
  }
  // the steps still queued (guesses past the last level) return at once
  b->vis.p = cur.vis;
  b->vis_alt.p = cur.vis_alt;
  b->q.p = cur.q;
  b->nq.p = cur.nq;
  // the gather: per vertex with edges its original ID, new ID, entry read and
  // written; with CoreBFS every original ID (new ID and entry written) and
  // every vertex with edges' entry
  int64_t gather_bytes = b->n_core > 0 ? 24 * b->n_core : 12 * g->n + 8 * n_c;
  const bool disjoint = early_gather && disjoint_fixup(b);
  if (early_gather) {
    // unless the guessed fixup followed the step that ended the run
    if (!(eg.fixup_seq == ran.back().seq && eg.fixup_code == final_code))
      queue_fixup(b, b->vis.get(), 0);
    // both bitmaps; per vertex its new -> original ID, entry read and written
    alg += 16 * words + 20 * after_snap;
    if (disjoint) {
      // the fixup also writes -1 for the vertices never reached, which the
      // gather leaves out with those reached after the snapshot
      const int64_t unreached = std::max<int64_t>(0, n_c - hl.visited);
      alg += 12 * unreached;
      gather_bytes -= 8 * (after_snap + unreached);
    }
  } else {
    launch_gather(b, root, s, 0, 0u);
  }
  alg += gather_bytes;
  if (b->n_core > 0) {
    // the gather skips vertices without edges: a root among them gets its
    // own entry here, and the last such root's entry goes back to -1
    if (b->iso_root >= 0 && b->iso_root != root) path.emplace_back((uint32_t)b->iso_root, -1);
    b->iso_root = -1;
    if (!root_in) {
      path.emplace_back((uint32_t)root, root);
      b->iso_root = root;
    }
  }
  for (size_t i = 0; i < path.size(); i += kMaxPathArgs) {
    PathArgs pa;
    pa.count = (int)std::min<size_t>(kMaxPathArgs, path.size() - i);
    for (int j = 0; j < pa.count; ++j) {
      pa.v[j] = path[i + j].first;
      pa.e[j] = path[i + j].second;
    }
    hipLaunchKernelGGL(k_write_path, dim3(1), dim3(kMaxPathArgs), 0, s, pa, b->pred.get());
    G500_CHECK_LAUNCH();
  }
  G500_CHECK_HIP(hipEventRecord(b->ev_end, s));
  G500_CHECK_HIP(hipEventSynchronize(b->ev_end));
  float ms = 0;
  G500_CHECK_HIP(hipEventElapsedTime(&ms, b->ev_begin, b->ev_end));
  if (disjoint) {
    // the gather may end after the engine stream's last kernel
    G500_CHECK_HIP(hipEventSynchronize(b->ev_g1));
    b->gather_pending = false;
    float gms = 0;
    G500_CHECK_HIP(hipEventElapsedTime(&gms, b->ev_begin, b->ev_g1));
    ms = std::max(ms, gms);
  }
  b->stats.gpu_ms = ms;
  b->stats.nlevels = (int)b->trace.size();
  b->stats.mispredicted = mispredicted;
  // vertices with edges this run left unreached: the next run's guesses
  // take them to lie in other components
  b->unreach = std::max<int64_t>(0, n_c - hl.visited);
  G500_CHECK_HIP(hipEventElapsedTime(&ms, b->ev_g0, b->ev_g1));
  b->stats.gather_ms = ms;
  for (const Step& st : ran) {
    if (st.sweep_ev < 0) continue;
    G500_CHECK_HIP(hipEventElapsedTime(&ms, b->sweep_events[st.sweep_ev],
                                       b->sweep_events[st.sweep_ev + 1]));
    b->stats.sweep_ms += ms;
    ++b->stats.sweeps;
  }
  if (b->counting) {
    std::vector<uint64_t> w(b->wacc.n);
    G500_CHECK_HIP(hipMemcpy(w.data(), b->wacc.get(), w.size() * sizeof(uint64_t),
                             hipMemcpyDeviceToHost));
    uint64_t sweep_dev = 0, other_dev = 0;
    for (size_t i = 0; i < w.size(); ++i) (i < b->bu_grid[0] ? sweep_dev : other_dev) += w[i];
    b->stats.counted = 1;
    b->stats.alg_bytes = alg + (int64_t)(sweep_dev + other_dev);
    b->stats.sweep_alg_bytes = b->stats.sweeps ? sweep_host + (int64_t)sweep_dev : 0;
    b->stats.gather_alg_bytes = gather_bytes;
  }
  if (b->timing) {
    // per level: its step's device time; per boundary: the time from the end
    // of one step that ran to the start of the next (host decisions, launch
    // latency and any guessed step that returned at once)
    size_t li_ = 0;
    for (size_t i = 0; i < ran.size(); ++i) {
      float lm = 0;
      G500_CHECK_HIP(hipEventElapsedTime(&lm, b->level_events[ran[i].ev],
                                         b->level_events[ran[i].ev + 1]));
      if (step_mode(ran[i].code) != kModeConv && li_ < b->trace.size()) {
        b->trace[li_].gpu_ms = lm;
        (b->trace[li_].top_down ? b->stats.td_ms : b->stats.bu_ms) += lm;
        ++li_;
      }
      if (i > 0) {
        G500_CHECK_HIP(hipEventElapsedTime(&lm, b->level_events[ran[i - 1].ev + 1],
                                           b->level_events[ran[i].ev]));
        b->stats.gap_ms += lm;
        ++b->stats.gaps;
      }
    }
  }
  for (const auto& tr : b->trace)
    if (tr.top_down) b->stats.td_edges += tr.nq_edges;
  b->stats.root_level = root_level;
  b->have_result = true;
  b->last_root = root;
}

void bfs_levels(g500_bfs* b, uint16_t* out) {
  g500_graph* g = b->g;
  DevBuf<uint16_t> lv;
  lv.alloc((size_t)g->n);
  hipLaunchKernelGGL(k_levels, dim3((unsigned)ceil_div(g->n, kT)), dim3(kT), 0, b->stream,
                     b->pred.get(), g->n, lv.get());
  G500_CHECK_LAUNCH();
  G500_CHECK_HIP(hipMemcpyAsync(out, lv.get(), g->n * sizeof(uint16_t),
                                hipMemcpyDeviceToHost, b->stream));
  G500_CHECK_HIP(hipStreamSynchronize(b->stream));
}

}  // namespace g500

g500_bfs::~g500_bfs() {
  if (h_report) (void)hipHostFree(h_report);
  if (ev_begin) (void)hipEventDestroy(ev_begin);
  if (ev_end) (void)hipEventDestroy(ev_end);
  if (ev_g0) (void)hipEventDestroy(ev_g0);
  if (ev_g1) (void)hipEventDestroy(ev_g1);
  for (auto e : sweep_events) (void)hipEventDestroy(e);
  if (gstream) (void)hipStreamDestroy(gstream);
  for (auto e : level_events) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
  if (ev_snap) (void)hipEventDestroy(ev_snap);
}
