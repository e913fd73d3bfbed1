// bfs.hip -- direction-optimizing BFS on one MI355X.
//
// Level loop and alpha/beta switch follow BfsBase::run_bfs_core
// (mpi/bfs.hpp:3238-3566, decision at 3413-3457).  The data layout is our
// own: vertices renumbered by descending degree ("new" IDs), a uint64 visited
// bitmap over new IDs, uint32 frontier queues, and pred written directly by
// original ID as packed (parent | depth << 48) entries.
//
// Kernels
//   k_td_expand   top-down: one prefix sum over frontier degrees, then a
//                 persistent grid walks 2048-edge tiles; the tile's slice of
//                 the prefix sum sits in LDS and each lane binary-searches
//                 its edge's owner there (every lane gets one edge, whatever
//                 the degree skew).  Winners of the visited-bit atomicOr write
//                 pred and are appended to an LDS queue that is flushed with
//                 one global atomic per tile.
//   k_bu_step     bottom-up: a wave reads 64 consecutive bitmap words
//                 (coalesced), compacts the unvisited vertices onto its 64
//                 lanes with shuffles, and each lane scans its vertex's row
//                 until the first visited neighbour (early exit).  New bits
//                 gather in LDS and are written once per word.
//   k_bitmap_to_queue   bitmap -> queue at the bottom-up -> top-down switch.
#include <algorithm>
#include <cstring>

#include "engine.hpp"

namespace g500 {

namespace {

constexpr int kT = 256;
constexpr int kWavesPerBlock = kT / kWave;
constexpr int kTdItems = 8;
constexpr int kTdTile = kT * kTdItems;
constexpr int kMaxGrid = 2048;

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

// Sum over the block; valid in thread 0.
__device__ inline uint64_t block_sum(uint64_t v, uint64_t* lds) {
  v = wave_sum_u64(v);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  uint64_t s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWavesPerBlock; ++w) s += lds[w];
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

// Bits of vertices >= n_c are pre-set so they are never searched
// (the reference pre-visits vertices without edges, mpi/bfs.hpp:330-334).
__global__ __launch_bounds__(kT) void k_bfs_init(uint64_t* __restrict__ vis,
                                                int64_t words, int64_t n_c,
                                                int64_t root_new) {
  for (int64_t w = (int64_t)blockIdx.x * kT + threadIdx.x; w < words;
       w += (int64_t)gridDim.x * kT) {
    const int64_t lo = w * 64;
    const int64_t valid = n_c - lo;
    uint64_t word = 0;
    if (valid < 64) word = valid <= 0 ? ~0ull : (~0ull << valid);
    if (root_new >= lo && root_new < lo + 64) word |= 1ull << (root_new - lo);
    vis[w] = word;
  }
}

__global__ void k_bfs_root(int64_t* pred, int64_t root, int64_t entry,
                           uint32_t* q, int64_t root_new, int64_t n_c) {
  pred[root] = entry;
  if (root_new >= 0 && root_new < n_c) q[0] = (uint32_t)root_new;
}

__global__ __launch_bounds__(kT) void k_td_degrees(const uint32_t* __restrict__ q,
                                                  int64_t qlen,
                                                  const int64_t* __restrict__ row_off,
                                                  uint64_t* __restrict__ dsum) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i < qlen) {
    const uint32_t u = q[i];
    dsum[i] = (uint64_t)(row_off[u + 1] - row_off[u]);
  }
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

__global__ __launch_bounds__(kT) void k_td_expand(
    const uint32_t* __restrict__ q, int64_t qlen,
    const uint64_t* __restrict__ dsum, const int64_t* __restrict__ row_off,
    const uint32_t* __restrict__ col, const uint32_t* __restrict__ n2o,
    const uint32_t* __restrict__ deg_new, unsigned long long* vis,
    int64_t* __restrict__ pred, uint32_t* __restrict__ nq,
    unsigned long long* counters, uint64_t level_bits) {
  __shared__ uint64_t win[kTdTile];
  __shared__ uint32_t lq[kTdTile];
  __shared__ uint64_t red[kWavesPerBlock];
  __shared__ int64_t s_lo, s_hi;
  __shared__ uint32_t lq_n;
  __shared__ unsigned long long lq_base;

  const int t = threadIdx.x;
  const uint64_t E = dsum[qlen];
  uint64_t my_edges = 0;
  for (uint64_t e0 = (uint64_t)blockIdx.x * kTdTile; e0 < E;
       e0 += (uint64_t)gridDim.x * kTdTile) {
    const uint64_t e1 = (e0 + kTdTile < E) ? e0 + kTdTile : E;
    if (t == 0) {
      const int64_t lo = owner_search(dsum, 0, qlen - 1, e0);
      s_lo = lo;
      s_hi = owner_search(dsum, lo, qlen - 1, e1 - 1);
      lq_n = 0;
    }
    __syncthreads();
    const int64_t lo = s_lo, hi = s_hi;
    const int64_t wn = hi - lo + 1;
    const bool use_win = wn <= kTdTile;
    if (use_win)
      for (int64_t j = t; j < wn; j += kT) win[j] = dsum[lo + j];
    __syncthreads();
    for (int it = 0; it < kTdItems; ++it) {
      const uint64_t e = e0 + (uint64_t)it * kT + t;
      bool won = false;
      uint32_t v = 0;
      if (e < e1) {
        int64_t i;
        uint64_t base;
        if (use_win) {
          const int64_t li = owner_search(win, 0, wn - 1, e);
          i = lo + li;
          base = win[li];
        } else {
          i = owner_search(dsum, lo, hi, e);
          base = dsum[i];
        }
        const uint32_t u = q[i];
        v = col[row_off[u] + (int64_t)(e - base)];
        const unsigned long long bit = 1ull << (v & 63);
        unsigned long long* wp = vis + (v >> 6);
        if (!(__atomic_load_n(wp, __ATOMIC_RELAXED) & bit)) {
          const unsigned long long old = atomicOr(wp, bit);
          won = !(old & bit);
        }
        if (won) {
          pred[n2o[v]] = (int64_t)((uint64_t)n2o[u] | level_bits);
          my_edges += deg_new[v];
        }
      }
      const uint64_t mask = __ballot(won);
      if (mask) {
        uint32_t wbase = 0;
        if ((mask & lanemask_lt()) == 0 && won)
          wbase = atomicAdd(&lq_n, (uint32_t)__popcll(mask));
        const int leader = __ffsll((long long)mask) - 1;
        wbase = __shfl(wbase, leader, kWave);
        if (won) lq[wbase + __popcll(mask & lanemask_lt())] = v;
      }
    }
    __syncthreads();
    const uint32_t cnt = lq_n;
    if (t == 0 && cnt) lq_base = atomicAdd(&counters[0], (unsigned long long)cnt);
    __syncthreads();
    for (uint32_t j = t; j < cnt; j += kT) nq[lq_base + j] = lq[j];
    __syncthreads();
  }
  const uint64_t s = block_sum(my_edges, red);
  if (t == 0 && s) atomicAdd(&counters[1], (unsigned long long)s);
}

__global__ __launch_bounds__(kT) void k_bu_step(
    const uint64_t* __restrict__ vis_old, uint64_t* __restrict__ vis_new,
    int64_t words, const int64_t* __restrict__ row_off,
    const uint32_t* __restrict__ col, const uint32_t* __restrict__ n2o,
    int64_t* __restrict__ pred, unsigned long long* counters,
    uint64_t level_bits) {
  __shared__ unsigned long long fw[kWavesPerBlock][kWave];
  __shared__ uint64_t red[kWavesPerBlock];
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  const int64_t nchunks = ceil_div(words, kWave);
  uint64_t nfound = 0;
  for (int64_t c = (int64_t)blockIdx.x * kWavesPerBlock + wave; c < nchunks;
       c += (int64_t)gridDim.x * kWavesPerBlock) {
    const int64_t w = c * kWave + lane;
    const uint64_t word = w < words ? vis_old[w] : ~0ull;
    const uint64_t um = ~word;
    const uint32_t cnt = (uint32_t)__popcll(um);
    uint32_t incl = cnt;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, kWave);
      if (lane >= o) incl += y;
    }
    const uint32_t excl = incl - cnt;
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    fw[wave][lane] = 0;
    wave_sync();
    for (uint32_t b0 = 0; b0 < total; b0 += kWave) {
      const uint32_t gidx = b0 + lane;
      // owning word: largest l with excl_l <= gidx
      int lo = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const int cand = lo + step;
        const uint32_t e = __shfl(excl, cand < kWave ? cand : kWave - 1, kWave);
        if (cand < kWave && e <= gidx) lo = cand;
      }
      const uint64_t wum = __shfl(um, lo, kWave);
      const uint32_t elo = __shfl(excl, lo, kWave);
      if (gidx < total) {
        uint32_t r = gidx - elo;
        uint64_t x = wum;
        int pos = 0;
#pragma unroll
        for (int half = 32; half > 0; half >>= 1) {
          const uint32_t lc = (uint32_t)__popcll(x & ((1ull << half) - 1));
          if (r >= lc) {
            r -= lc;
            x >>= half;
            pos += half;
          }
        }
        const uint32_t v = (uint32_t)((c * kWave + lo) * 64 + pos);
        const int64_t k0 = row_off[v], k1 = row_off[v + 1];
        for (int64_t k = k0; k < k1; ++k) {
          const uint32_t u = col[k];
          if ((vis_old[u >> 6] >> (u & 63)) & 1ull) {
            atomicOr(&fw[wave][lo], 1ull << pos);
            pred[n2o[v]] = (int64_t)((uint64_t)n2o[u] | level_bits);
            break;
          }
        }
      }
    }
    wave_sync();
    const uint64_t fbits = fw[wave][lane];
    if (w < words) vis_new[w] = word | fbits;
    nfound += (uint64_t)__popcll(fbits);
    wave_sync();
  }
  const uint64_t s = block_sum(nfound, red);
  if (threadIdx.x == 0 && s) atomicAdd(&counters[0], (unsigned long long)s);
}

// queue <- bits set in cur but not in prev; counters[2] += count,
// counters[3] += degree sum.
__global__ __launch_bounds__(kT) void k_bitmap_to_queue(
    const uint64_t* __restrict__ cur, const uint64_t* __restrict__ prev,
    int64_t words, const uint32_t* __restrict__ deg_new,
    uint32_t* __restrict__ out, unsigned long long* counters) {
  __shared__ uint64_t wsum[kWavesPerBlock];
  __shared__ unsigned long long base;
  const int64_t w = (int64_t)blockIdx.x * kT + threadIdx.x;
  uint64_t bits = w < words ? (cur[w] & ~prev[w]) : 0;
  uint64_t tot;
  const uint64_t excl = block_excl_scan((uint64_t)__popcll(bits), wsum, &tot);
  if (threadIdx.x == 0) base = tot ? atomicAdd(&counters[2], (unsigned long long)tot) : 0;
  __syncthreads();
  uint64_t pos = base + excl, edges = 0;
  while (bits) {
    const int b = __ffsll((long long)bits) - 1;
    bits &= bits - 1;
    const uint32_t v = (uint32_t)(w * 64 + b);
    out[pos++] = v;
    edges += deg_new[v];
  }
  const uint64_t s = block_sum(edges, wsum);
  if (threadIdx.x == 0 && s) atomicAdd(&counters[3], (unsigned long long)s);
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

void bfs_init(g500_bfs* b) {
  g500_graph* g = b->g;
  b->words = ceil_div(std::max<int64_t>(g->n_c, 1), 64);
  b->pred.alloc((size_t)g->n);
  b->vis.alloc((size_t)b->words);
  b->vis_alt.alloc((size_t)b->words);
  b->q.alloc((size_t)std::max<int64_t>(g->n_c, 1));
  b->nq.alloc((size_t)std::max<int64_t>(g->n_c, 1));
  b->dsum.alloc((size_t)std::max<int64_t>(g->n_c, 1) + 1);
  b->counters.alloc(8);
  b->scan_ws.reserve(std::max<int64_t>(g->n_c, 1));
  G500_CHECK_HIP(hipHostMalloc((void**)&b->h_counters, 8 * sizeof(unsigned long long),
                               hipHostMallocDefault));
  G500_CHECK_HIP(hipEventCreate(&b->ev_begin));
  G500_CHECK_HIP(hipEventCreate(&b->ev_end));
  ensure_host_orig2new(g);
}

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
  const int64_t root_new = g->h_orig2new.empty() ? -1 : (int64_t)g->h_orig2new[root];
  const bool root_in = root_new >= 0 && root_new < n_c;

  b->trace.clear();
  b->stats = g500_run_stats{};
  b->have_result = false;
  G500_CHECK_HIP(hipEventRecord(b->ev_begin, s));
  G500_CHECK_HIP(hipMemsetAsync(b->pred.get(), 0xFF, g->n * sizeof(int64_t), s));
  hipLaunchKernelGGL(k_bfs_init, dim3(grid_for(words, kT)), dim3(kT), 0, s,
                     b->vis.get(), words, n_c, root_in ? root_new : (int64_t)-1);
  G500_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_bfs_root, dim3(1), dim3(1), 0, s, b->pred.get(), root,
                     root /* depth 0 */, b->q.get(), root_in ? root_new : (int64_t)-1,
                     n_c);
  G500_CHECK_LAUNCH();

  int64_t qlen = root_in ? 1 : 0;
  int64_t est_edges = 0;
  if (root_in) {
    int64_t ro[2];
    G500_CHECK_HIP(hipMemcpyAsync(ro, g->row_off.get() + root_new, 2 * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
    est_edges = ro[1] - ro[0];
  }

  // level-loop state (mpi/bfs.hpp:3264-3279)
  bool next_td = true, growing = true;
  int64_t prev_nq = 0;
  int level = 0;
  const int64_t alpha_thr = (int64_t)((double)g->m / alpha);
  const int64_t beta_thr = (int64_t)((double)n_c / (beta * edgefactor * 2.0));
  unsigned long long* cnt = b->counters.get();

  for (;;) {
    at[G500_AT_LEVEL] = ++level;
    const bool td = next_td;
    hipEvent_t lev0 = nullptr, lev1 = nullptr;
    if (b->timing) {
      const size_t k = b->trace.size() * 2;
      while (b->level_events.size() < k + 2) {
        hipEvent_t e;
        G500_CHECK_HIP(hipEventCreate(&e));
        b->level_events.push_back(e);
      }
      lev0 = b->level_events[k];
      lev1 = b->level_events[k + 1];
      G500_CHECK_HIP(hipEventRecord(lev0, s));
    }
    G500_CHECK_HIP(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned long long), s));
    const uint64_t level_bits = (uint64_t)level << 48;
    if (td) {
      if (qlen > 0) {
        hipLaunchKernelGGL(k_td_degrees, dim3((unsigned)ceil_div(qlen, kT)), dim3(kT),
                           0, s, b->q.get(), qlen, g->row_off.get(), b->dsum.get());
        G500_CHECK_LAUNCH();
      }
      exclusive_scan<uint64_t>(b->dsum.get(), b->dsum.get(), qlen, b->scan_ws, s);
      hipLaunchKernelGGL(k_td_expand, dim3(grid_for(est_edges, kTdTile)), dim3(kT), 0, s,
                         b->q.get(), qlen, b->dsum.get(), g->row_off.get(),
                         g->col.get(), g->new2orig.get(), g->deg_new.get(),
                         (unsigned long long*)b->vis.get(), b->pred.get(), b->nq.get(),
                         cnt, level_bits);
      G500_CHECK_LAUNCH();
    } else {
      hipLaunchKernelGGL(k_bu_step, dim3(grid_for(ceil_div(words, kWave), kWavesPerBlock)),
                         dim3(kT), 0, s, b->vis.get(), b->vis_alt.get(), words,
                         g->row_off.get(), g->col.get(), g->new2orig.get(),
                         b->pred.get(), cnt, level_bits);
      G500_CHECK_LAUNCH();
      std::swap(b->vis.p, b->vis_alt.p);
    }
    if (b->timing) G500_CHECK_HIP(hipEventRecord(lev1, s));
    G500_CHECK_HIP(hipMemcpyAsync(b->h_counters, cnt, 4 * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
    const int64_t nq_size = (int64_t)b->h_counters[0];
    const int64_t nq_edges = td ? (int64_t)b->h_counters[1] : 0;
    g500_level_trace tr{};
    tr.level = level;
    tr.top_down = td ? 1 : 0;
    tr.nq_size = nq_size;
    tr.nq_edges = nq_edges;
    if (b->timing) {
      float ms = 0;
      G500_CHECK_HIP(hipEventElapsedTime(&ms, lev0, lev1));
      tr.gpu_ms = ms;
      (td ? b->stats.td_ms : b->stats.bu_ms) += ms;
    }
    b->trace.push_back(tr);
    if (nq_size == 0) break;

    // direction decision (mpi/bfs.hpp:3413-3457)
    if (growing && nq_size > prev_nq) {
      if (td) {
        if (nq_edges > alpha_thr) {
          next_td = false;
          at[G500_AT_GLOBAL_NQ_EDGES] = nq_edges;
          at[G500_AT_NUM_GLOBAL_EDGES] = g->m;
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

    if (td) {
      std::swap(b->q.p, b->nq.p);
      qlen = nq_size;
      est_edges = nq_edges;
    } else if (next_td) {
      // switching to top-down: frontier = bits new in this level
      G500_CHECK_HIP(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_bitmap_to_queue, dim3((unsigned)ceil_div(words, kT)), dim3(kT),
                         0, s, b->vis.get(), b->vis_alt.get(), words, g->deg_new.get(),
                         b->q.get(), cnt);
      G500_CHECK_LAUNCH();
      G500_CHECK_HIP(hipMemcpyAsync(b->h_counters, cnt, 4 * sizeof(unsigned long long),
                                    hipMemcpyDeviceToHost, s));
      G500_CHECK_HIP(hipStreamSynchronize(s));
      qlen = (int64_t)b->h_counters[2];
      est_edges = (int64_t)b->h_counters[3];
      if (qlen != nq_size)
        fail(G500_ERR_STATE, "frontier conversion mismatch (%lld vs %lld)",
             (long long)qlen, (long long)nq_size);
    }
    prev_nq = nq_size;
  }
  G500_CHECK_HIP(hipEventRecord(b->ev_end, s));
  G500_CHECK_HIP(hipEventSynchronize(b->ev_end));
  float ms = 0;
  G500_CHECK_HIP(hipEventElapsedTime(&ms, b->ev_begin, b->ev_end));
  b->stats.gpu_ms = ms;
  b->stats.nlevels = (int)b->trace.size();
  for (const auto& tr : b->trace)
    if (tr.top_down) b->stats.td_edges += tr.nq_edges;
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
  if (h_counters) (void)hipHostFree(h_counters);
  if (ev_begin) (void)hipEventDestroy(ev_begin);
  if (ev_end) (void)hipEventDestroy(ev_end);
  for (auto e : level_events) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}
