// msbfs.hip -- batched bit-parallel multi-source BFS on one MI355X.
//
// R <= 256 roots traverse the graph together.  Vertex v (new ID) carries
// K = ceil(R/64) lane words: seen[v] (lanes that reached v) and, per level l,
// F_l[v] (lanes whose level-l frontier holds v), plus one bit per vertex in a
// frontier bitmap (F_l[v] != 0).  Levels are unique per root, so the per-root
// level arrays equal R independent BFS runs whatever order the lanes are
// discovered in.  Each level goes one of two ways, chosen on the device from
// the frontier's summed degree fe[l] (the rule of a direction-optimizing BFS,
// with its own threshold m / td_alpha):
//
//   top-down   every frontier vertex pushes its lane word to its neighbours
//              (atomicOr into F_{l+1}, filtered by seen)
//   bottom-up  every vertex not yet seen by all lanes ORs F_l over those
//              neighbours whose frontier bit is set, until every lane it lacks
//              is covered (early exit); F_{l+1}[v] = acc & ~seen[v]
//
// Vertices are degree-descending by new ID, so both sweeps split [0, n_c)
// into three contiguous bins: heavy vertices cut into fixed work items of
// `heavy_chunk` edges (one wave each, partial pulls merged in an accumulator),
// medium vertices with a group of lanes each, light vertices with one lane
// each.  Every level kernel first reads the level's discovery counter and
// returns at once when it is zero, so a fixed number of levels is captured in
// one HIP graph and replayed without a host round trip per level.  After the
// traversal one kernel turns the level frontiers into per-root uint8 level
// arrays by original ID (bit-plane accumulation and an 8x8 bit transpose), and
// one sums per-root traversed edges.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "msbfs.hpp"

namespace g500 {
namespace {

constexpr int kT = 256;
constexpr int kWavesPerBlock = kT / kWave;
constexpr int kMaxLevels = 255;    // uint8 levels 0..254, 0xFF = unreached
// level buffers F_0..F_255: the last one only proves that depth 255 is empty
constexpr int kInitialLevels = 12;  // Kronecker graphs need 8-10; deeper ones grow

__device__ inline uint64_t pad_mask(int k, int nroots) {
  const int nvalid = min(max(nroots - 64 * k, 0), 64);
  return nvalid == 64 ? 0ull : ~((1ull << nvalid) - 1ull);
}

__device__ inline uint64_t block_sum(uint64_t v, uint64_t* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  if (lane_id() == 0) lds[threadIdx.x / kWave] = v;
  __syncthreads();
  uint64_t s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWavesPerBlock; ++w) s += lds[w];
  return s;
}

// Input edges by source vertex, by new ID: the traversed-edge count of a root
// is the sum over the vertices it reaches (the validator counts an input edge
// when its source is reached).
__global__ __launch_bounds__(kT) void k_ms_cnt0(const uint2* __restrict__ e, int64_t ne,
                                               const uint32_t* __restrict__ o2n,
                                               uint32_t* __restrict__ cnt0) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < ne;
       i += (int64_t)gridDim.x * kT)
    atomicAdd(&cnt0[o2n[e[i].x]], 1u);
}

// First new ID whose degree is below each threshold (deg_new is descending).
__global__ void k_ms_bins(const uint32_t* __restrict__ deg, int64_t n_c, int64_t heavy,
                          int64_t medium, int64_t* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t th[2] = {heavy, medium};
  for (int t = 0; t < 2; ++t) {
    int64_t lo = 0, hi = n_c;  // first index with deg < th
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if ((int64_t)deg[mid] >= th[t]) lo = mid + 1;
      else hi = mid;
    }
    out[t] = lo;
  }
}

// Per-batch reset: seen = padding lanes, F_0 = 0, frontier bitmaps, heavy
// accumulators, counters and TE = 0.  d_roots[0] holds the batch size,
// d_roots[1..] the roots.
__global__ __launch_bounds__(kT) void k_ms_init(const int64_t* __restrict__ d_roots,
                                               int64_t n_c, int K,
                                               uint64_t* __restrict__ seen,
                                               uint64_t* __restrict__ f0,
                                               uint64_t* __restrict__ fb0,
                                               uint64_t* __restrict__ fb1, int64_t bm_words,
                                               uint64_t* __restrict__ acc, int64_t acc_words,
                                               unsigned long long* __restrict__ cnt,
                                               unsigned long long* __restrict__ fe, int ncnt,
                                               unsigned long long* __restrict__ te,
                                               uint64_t* __restrict__ act) {
  const int nroots = (int)d_roots[0];
  const int64_t words = n_c * K;
  const int64_t i0 = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kT;
  for (int64_t i = i0; i < words; i += stride) {
    seen[i] = pad_mask((int)(i % K), nroots);
    f0[i] = 0;
  }
  for (int64_t i = i0; i < bm_words; i += stride) {
    fb0[i] = 0;
    fb1[i] = 0;
  }
  for (int64_t i = i0; i < acc_words; i += stride) acc[i] = 0;
  for (int64_t j = i0; j < ncnt; j += stride) {
    cnt[j] = 0;
    fe[j] = 0;
  }
  for (int64_t j = i0; j < (int64_t)ncnt * 4; j += stride) act[j] = 0;
  for (int64_t j = i0; j < G500_MSBFS_MAX_ROOTS; j += stride) te[j] = 0;
}

__global__ void k_ms_roots(const int64_t* __restrict__ d_roots,
                           const uint32_t* __restrict__ o2n,
                           const uint32_t* __restrict__ deg, int64_t n_c, int K,
                           uint64_t* __restrict__ seen, uint64_t* __restrict__ f0,
                           uint64_t* __restrict__ fb0, unsigned long long* cnt,
                           unsigned long long* fe, uint64_t* act) {
  const int nroots = (int)d_roots[0];
  const int i = threadIdx.x;
  if (i >= nroots) return;
  const int64_t v = o2n[d_roots[1 + i]];
  if (v >= n_c) return;  // no edges: handled by k_ms_isolated
  const uint64_t bit = 1ull << (i & 63);
  atomicOr((unsigned long long*)&f0[v * K + (i >> 6)], (unsigned long long)bit);
  atomicOr((unsigned long long*)&seen[v * K + (i >> 6)], (unsigned long long)bit);
  atomicOr((unsigned long long*)&act[i >> 6], (unsigned long long)bit);
  const unsigned long long vb = 1ull << (v & 63);
  if (!(atomicOr((unsigned long long*)&fb0[v >> 6], vb) & vb)) atomicAdd(&fe[0], (unsigned long long)deg[v]);
  atomicAdd(&cnt[0], 1ull);
}

__device__ inline uint32_t ld_col(const uint32_t* p) { return *p; }
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__device__ inline u32x4_t ld_col4(const uint32_t* p) {
  return *reinterpret_cast<const u32x4_t*>(p);
}

template <int K>
__device__ inline void load_lanes(const uint64_t* __restrict__ p, uint64_t (&x)[K]) {
  if constexpr (K == 2 || K == 4) {
    const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
#pragma unroll
    for (int j = 0; j < K / 2; ++j) {
      const ulonglong2 y = q[j];
      x[2 * j] = y.x;
      x[2 * j + 1] = y.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = p[k];
  }
}

// What every kernel of level L -> L+1 sees.  Each kernel first checks that
// level L has a frontier (else the traversal is over and it returns), then
// which direction the level takes: top-down while the frontier's summed
// degree fe[L] is below thr, else bottom-up.  Both are recomputed from device
// counters, so a captured level needs no host decision.
struct MsLevel {
  const int64_t* row_off;
  const uint32_t* col;
  const uint32_t* deg;
  uint64_t* seen;
  const uint64_t* fl;   // F_L
  uint64_t* fn;         // F_{L+1}
  const uint64_t* fb;   // bitmap of F_L
  uint64_t* fbn;        // bitmap of F_{L+1}
  unsigned long long* cnt;
  unsigned long long* fe;
  const uint64_t* act_l;   // lanes present anywhere in F_L (4 words)
  uint64_t* act_n;         // the same for F_{L+1}, ORed by the kernels that write it
  unsigned long long thr;
  const int64_t* d_roots;  // [0] = batch size
  double stream_frac;
  int64_t n_c;
  int64_t hot;             // dense_algo 2: vertices below this are pulled
  int dense_algo;          // dense levels: 1 streaming pull, 2 hub pull + push
  int L;
  // split level: roots whose frontier degree sum is below split_thr push;
  // lane_part holds per-block degree sums per lane, split_mask the lanes
  // that push (written by k_ms_lane_mask before the level's pulls)
  unsigned long long split_thr;  // 0: no split levels
  unsigned long long* lane_part;
  uint64_t* split_mask;
};

enum { kModeTopDown = 0, kModeStream = 1, kModePull = 2, kModeSplit = 3 };
// rows of per-lane degree sums for a split level: block b of a top-down
// level's count pass adds its sums to row b % kLaneParts
constexpr int kLaneParts = 512;

// Adds degree d to the LDS counter of every lane set in x.  A word with more
// lanes set than clear adds d to the all-lanes counter and subtracts it from
// its clear lanes (modulo 2^64), so no word costs more than 33 LDS atomics.
__device__ inline void lane_add(unsigned long long* lfe, unsigned long long* lall, uint64_t x,
                                unsigned long long d) {
  if (__popcll(x) > 32) {
    atomicAdd(lall, d);
    for (uint64_t b = ~x; b; b &= b - 1) atomicAdd(&lfe[__ffsll((long long)b) - 1], 0ull - d);
  } else {
    for (uint64_t b = x; b; b &= b - 1) atomicAdd(&lfe[__ffsll((long long)b) - 1], d);
  }
}
enum { kDenseStream = 1, kDenseHybrid = 2 };

__device__ inline bool ms_active(const MsLevel& a) { return a.cnt[a.L] != 0; }
__device__ inline bool ms_top_down(const MsLevel& a, int l) { return a.fe[l] < a.thr; }
// top-down while the frontier is light; else a streaming pull while few
// (root, vertex) pairs are reached yet, then a per-vertex pull -- split at
// the first bottom-up level after top-down ones (light roots push)
__device__ inline int ms_mode(const MsLevel& a, int l) {
  if (ms_top_down(a, l)) return kModeTopDown;
  unsigned long long reached = 0;
  for (int j = 0; j <= l; ++j) reached += a.cnt[j];
  const double pairs = (double)a.d_roots[0] * (double)a.n_c;
  if ((double)reached < a.stream_frac * pairs) return kModeStream;
  return (a.split_thr > 0 && l > 0 && ms_top_down(a, l - 1)) ? kModeSplit : kModePull;
}
__device__ inline bool in_bitmap(const uint64_t* bm, int64_t v) {
  return (bm[v >> 6] >> (v & 63)) & 1ull;
}

__device__ inline void ms_flush(const MsLevel& a, uint64_t* red, uint64_t disc, uint64_t edges) {
  const uint64_t d = block_sum(disc, red);
  __syncthreads();
  const uint64_t e = block_sum(edges, red);
  if (threadIdx.x == 0) {
    if (d) atomicAdd(&a.cnt[a.L + 1], (unsigned long long)d);
    if (e) atomicAdd(&a.fe[a.L + 1], (unsigned long long)e);
  }
}

// ORs the lanes a thread wrote to F_{L+1} into act_n: a wave reduction, then
// one atomic per word per wave, skipped when the word already holds them.
template <int K>
__device__ inline void ms_note_lanes(const MsLevel& a, uint64_t (&lanes)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lanes[k] |= __shfl_xor(lanes[k], o, kWave);
    if (lane_id() == 0 && lanes[k]) {
      const uint64_t have = __hip_atomic_load(&a.act_n[k], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
      if (lanes[k] & ~have) atomicOr((unsigned long long*)&a.act_n[k], (unsigned long long)lanes[k]);
    }
  }
}

// Level preparation: clear the next bitmap; unless the level only pulls,
// clear F_{L+1} (pushes fill it by atomicOr; a pulling level stores every
// vertex's word); when F_L came from a top-down level and this level pushes
// too, fold F_L into seen (a level that pulls ORs F_L into each vertex's
// seen word itself, and the other modes write seen themselves); before a
// top-down level, clear the per-lane degree sums its count pass adds up for
// the next level.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_prep(MsLevel a, int64_t bm_words) {
  if (!ms_active(a)) return;
  const int mode = ms_mode(a, a.L);
  const bool clear = mode != kModePull;
  const bool fold = a.L > 0 && ms_top_down(a, a.L - 1) && mode == kModeTopDown;
  const int64_t i0 = (int64_t)blockIdx.x * kT + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * kT;
  for (int64_t i = i0; i < bm_words; i += stride) a.fbn[i] = 0;
  if (mode == kModeTopDown && a.split_thr > 0)
    for (int64_t i = i0; i < (int64_t)kLaneParts * 256; i += stride) a.lane_part[i] = 0;
  if (clear)
    for (int64_t i = i0; i < a.n_c * K; i += stride) a.fn[i] = 0;
  if (fold)
    for (int64_t w = i0; w < bm_words; w += stride)
      for (uint64_t x = a.fb[w]; x; x &= x - 1) {
        const int64_t v = w * 64 + __ffsll((long long)x) - 1;
#pragma unroll
        for (int k = 0; k < K; ++k) a.seen[v * K + k] |= a.fl[v * K + k];
      }
}

// Pushes run kPushIlp edges per lane at a time: the targets' seen words and
// bitmap words are all requested before the first atomic, since a wave
// waits for its atomics along with its loads (one vector memory counter),
// so a loop of single pushes pays three round trips per edge.
constexpr int kPushIlp = 4;

// The row entries [e0 + u*step] (u < kPushIlp) below end: targets, and
// whether each is in the row.
__device__ inline void push_targets(const uint32_t* col, int64_t e0, int64_t step, int64_t end,
                                    uint32_t (&w)[kPushIlp], bool (&ok)[kPushIlp]) {
#pragma unroll
  for (int u = 0; u < kPushIlp; ++u) {
    const int64_t e = e0 + (int64_t)u * step;
    ok[u] = e < end;
    w[u] = ok[u] ? ld_col(col + e) : 0u;
  }
}

// Top-down: push f to the targets w (lanes not yet seen there; a target
// with !ok pushes nothing).  The ORs return nothing, so a push waits on no
// atomic: the level's counts (new pairs, frontier degree sum) are taken
// afterwards from F_{L+1} by k_ms_td_count.  The bitmap word is read beside
// seen so a vertex already in the next frontier costs no second atomic.
template <int K>
__device__ inline void ms_push_n(const MsLevel& a, const uint32_t (&w)[kPushIlp],
                                 const bool (&ok)[kPushIlp], const uint64_t (&f)[K]) {
  uint64_t sw[kPushIlp][K], bw[kPushIlp];
#pragma unroll
  for (int u = 0; u < kPushIlp; ++u) {
    load_lanes<K>(a.seen + (int64_t)w[u] * K, sw[u]);
    bw[u] = __hip_atomic_load(&a.fbn[w[u] >> 6], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int u = 0; u < kPushIlp; ++u) {
    const uint64_t bit = 1ull << (w[u] & 63);
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t m = ok[u] ? f[k] & ~sw[u][k] : 0ull;
      if (m) {
        atomicOr((unsigned long long*)&a.fn[(int64_t)w[u] * K + k], (unsigned long long)m);
        any = true;
      }
    }
    if (any && !(bw[u] & bit))
      atomicOr((unsigned long long*)&a.fbn[w[u] >> 6], (unsigned long long)bit);
  }
}

// Pushes without a seen filter (the pulls mask F_{L+1} with seen): the
// lanes of f missing from each target's word, ORed in; lanes only ever get
// added, so a stale read can only cause an extra atomic.
template <int K>
__device__ inline void hyb_push_n(const MsLevel& a, const uint32_t (&w)[kPushIlp],
                                  const bool (&ok)[kPushIlp], const uint64_t (&f)[K]) {
  uint64_t x[kPushIlp][K];
#pragma unroll
  for (int u = 0; u < kPushIlp; ++u)
#pragma unroll
    for (int k = 0; k < K; ++k)
      x[u][k] = (uint64_t)__atomic_load_n((unsigned long long*)&a.fn[(int64_t)w[u] * K + k],
                                          __ATOMIC_RELAXED);
#pragma unroll
  for (int u = 0; u < kPushIlp; ++u)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t m = ok[u] ? f[k] & ~x[u][k] : 0ull;
      if (m) atomicOr((unsigned long long*)&a.fn[(int64_t)w[u] * K + k], (unsigned long long)m);
    }
}

// Pushes from the frontier F_L over new IDs [vb, ve), G lanes per vertex: at
// a top-down level every lane (filtered by seen: ms_push_n), at a split level
// only the lanes in split_mask (hyb_push_n: the pulls queued after them mask
// F_{L+1} with seen).  One kernel for both, so a level queues no push kernel
// that only returns.
template <int K>
__device__ inline bool push_lanes(const MsLevel& a, bool split, uint64_t (&pm)[K]) {
  bool anyp = false;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    pm[k] = split ? a.split_mask[k] : ~0ull;
    anyp = anyp || pm[k] != 0;
  }
  return anyp;
}

template <int K, int G>
__global__ __launch_bounds__(kT) void k_ms_push(MsLevel a, int64_t vb, int64_t ve) {
  if (!ms_active(a)) return;
  const int mode = ms_mode(a, a.L);
  if (mode != kModeTopDown && mode != kModeSplit) return;
  const bool split = mode == kModeSplit;
  uint64_t pm[K];
  if (!push_lanes<K>(a, split, pm)) return;
  const int lig = threadIdx.x % G;
  const int64_t gid = ((int64_t)blockIdx.x * kT + threadIdx.x) / G;
  const int64_t ngroups = (int64_t)gridDim.x * (kT / G);
  for (int64_t v = vb + gid; v < ve; v += ngroups) {
    if (!in_bitmap(a.fb, v)) continue;
    uint64_t f[K];
    load_lanes<K>(a.fl + v * K, f);
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      f[k] &= pm[k];
      any = any || f[k] != 0;
    }
    if (!any) continue;
    const int64_t end = a.row_off[v + 1];
    for (int64_t e0 = a.row_off[v] + lig; e0 < end; e0 += (int64_t)G * kPushIlp) {
      uint32_t w[kPushIlp];
      bool ok[kPushIlp];
      push_targets(a.col, e0, G, end, w, ok);
      if (split)
        hyb_push_n<K>(a, w, ok, f);
      else
        ms_push_n<K>(a, w, ok, f);
    }
  }
}

// The same for heavy vertices: one wave per work item of `chunk` edges.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_push_heavy(MsLevel a,
                                                     const uint64_t* __restrict__ item_start,
                                                     const uint32_t* __restrict__ item_vertex,
                                                     int64_t n_items, int chunk) {
  if (!ms_active(a)) return;
  const int mode = ms_mode(a, a.L);
  if (mode != kModeTopDown && mode != kModeSplit) return;
  const bool split = mode == kModeSplit;
  uint64_t pm[K];
  if (!push_lanes<K>(a, split, pm)) return;
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t it = wid; it < n_items; it += nw) {
    const int64_t v = item_vertex[it];
    if (!in_bitmap(a.fb, v)) continue;
    uint64_t f[K];
    load_lanes<K>(a.fl + v * K, f);
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      f[k] &= pm[k];
      any = any || f[k] != 0;
    }
    if (!any) continue;
    const int64_t beg = a.row_off[v] + (int64_t)(it - item_start[v]) * chunk;
    const int64_t end = min(a.row_off[v + 1], beg + (int64_t)chunk);
    for (int64_t e0 = beg + lane; e0 < end; e0 += (int64_t)kWave * kPushIlp) {
      uint32_t w[kPushIlp];
      bool ok[kPushIlp];
      push_targets(a.col, e0, kWave, end, w, ok);
      if (split)
        hyb_push_n<K>(a, w, ok, f);
      else
        ms_push_n<K>(a, w, ok, f);
    }
  }
}

// After a top-down level's pushes: new (vertex, root) pairs = the lanes of
// F_{L+1} (pushes left out lanes already seen), frontier degree sum = the
// degrees of the vertices in its bitmap.  One thread per vertex.
// With split levels on, it also adds up each lane's degree sum over F_{L+1}
// (the next level, if it is a split level, picks its pushing lanes from
// them): LDS counters per block, added to the block's lane_part row.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_td_count(MsLevel a) {
  __shared__ uint64_t red[kWavesPerBlock];
  __shared__ unsigned long long lfe[64 * K], lall[K];
  if (!ms_active(a) || !ms_top_down(a, a.L)) return;
  const bool split = a.split_thr > 0;
  if (split) {
    for (int i = threadIdx.x; i < 64 * K; i += kT) lfe[i] = 0;
    if (threadIdx.x < K) lall[threadIdx.x] = 0;
    __syncthreads();
  }
  uint64_t disc = 0, edges = 0, lanes[K];
#pragma unroll
  for (int k = 0; k < K; ++k) lanes[k] = 0;
  for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < a.n_c;
       v += (int64_t)gridDim.x * kT) {
    if (!in_bitmap(a.fbn, v)) continue;
    uint64_t x[K];
    load_lanes<K>(a.fn + v * K, x);
    const unsigned long long d = a.deg[v];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      disc += (uint64_t)__popcll(x[k]);
      lanes[k] |= x[k];
      if (split) lane_add(lfe + 64 * k, lall + k, x[k], d);
    }
    edges += d;
  }
  if (split) {
    __syncthreads();
    unsigned long long* row = a.lane_part + (int64_t)(blockIdx.x % kLaneParts) * 256;
    for (int i = threadIdx.x; i < 64 * K; i += kT) {
      const unsigned long long x = lfe[i] + lall[i / 64];
      if (x) atomicAdd(&row[i], x);
    }
  }
  ms_note_lanes<K>(a, lanes);
  ms_flush(a, red, disc, edges);
}

// Streaming pull: the edge array is cut into tiles of kTile edges, one block
// per tile, kTile / kT consecutive edges per thread (vector loads).  Each
// thread ORs the frontier lanes of its edges' targets per source vertex into
// LDS accumulators of the tile's vertex window; the window then goes to
// F_{L+1} (stored for rows inside the tile, atomicOr for the rows that cross
// its ends).  k_ms_commit masks the result with seen afterwards.
// Tile bounds of the streaming pull come from a per-graph table: entry t is
// the vertex whose row holds edge t * kStreamUnit.
constexpr int kStreamUnit = 1024;

template <int K>
struct StreamTile {
  static constexpr int kTile = K <= 2 ? 2048 : 1024;
  static constexpr int kEpt = kTile / kT;
};

template <int K>
__device__ inline void stream_flush(unsigned long long* acc_l, int64_t li, uint64_t (&run)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (run[k]) {
      atomicOr(&acc_l[li * K + k], (unsigned long long)run[k]);
      run[k] = 0;
    }
}

template <int K>
__global__ __launch_bounds__(kT) void k_ms_bu_stream(MsLevel a, int64_t m,
                                                    const uint32_t* __restrict__ tile_v0) {
  constexpr int kTile = StreamTile<K>::kTile, kEpt = StreamTile<K>::kEpt;
  constexpr int kStep = kTile / kStreamUnit;  // table entries per tile
  __shared__ int64_t win[kTile + 2];
  __shared__ unsigned long long acc_l[(kTile + 1) * K];
  __shared__ uint8_t full_l[kTile + 1];
  __shared__ int64_t s_v0, s_wn;
  if (!ms_active(a) || a.dense_algo != kDenseStream || ms_mode(a, a.L) != kModeStream) return;
  const int t = threadIdx.x;
  const int64_t ntiles = ceil_div(m, kTile);
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t e0 = tile * kTile, e1 = min(m, e0 + kTile);
    if (t == 0) {
      // owner of the tile's first edge, and of the next tile's first edge
      // (>= the owner of its last); rows are non-empty, so the window holds
      // at most kTile + 1 vertices
      const int64_t v0 = tile_v0[tile * kStep];
      const int64_t v1 = min((int64_t)tile_v0[(tile + 1) * kStep], a.n_c - 1);
      s_v0 = v0;
      s_wn = v1 - v0 + 1;
    }
    __syncthreads();
    const int64_t v0 = s_v0, wn = s_wn;
    for (int64_t i = t; i <= wn; i += kT) win[i] = a.row_off[v0 + i];
    for (int64_t i = t; i < wn * K; i += kT) acc_l[i] = 0;
    for (int64_t i = t; i < wn; i += kT) {
      uint64_t s0[K], f[K];
      load_lanes<K>(a.seen + (v0 + i) * K, s0);
      load_lanes<K>(a.fl + (v0 + i) * K, f);
      bool full = true;
#pragma unroll
      for (int k = 0; k < K; ++k) full = full && ((s0[k] | f[k] | ~a.act_l[k]) == ~0ull);
      full_l[i] = full ? 1 : 0;
    }
    __syncthreads();
    const int64_t eb = e0 + (int64_t)t * kEpt;
    if (eb < e1) {
      // phase 1: targets (vector loads) and owners (one search, then advance)
      uint32_t u[kEpt];
      if (eb + kEpt <= e1) {
#pragma unroll
        for (int j = 0; j < kEpt / 4; ++j) {
          const u32x4_t y = ld_col4(a.col + eb + 4 * j);
          u[4 * j] = y.x;
          u[4 * j + 1] = y.y;
          u[4 * j + 2] = y.z;
          u[4 * j + 3] = y.w;
        }
      } else {
#pragma unroll
        for (int j = 0; j < kEpt; ++j) u[j] = eb + j < e1 ? ld_col(a.col + eb + j) : 0u;
      }
      int64_t lo = 0, hi = wn - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (win[mid] <= eb) lo = mid;
        else hi = mid - 1;
      }
      int li[kEpt];
      {
        int cur = (int)lo;
#pragma unroll
        for (int j = 0; j < kEpt; ++j) {
          while (cur + 1 < wn && win[cur + 1] <= eb + j) ++cur;
          li[j] = cur;
        }
      }
      // phase 2: frontier bits of all targets
      bool hit[kEpt];
#pragma unroll
      for (int j = 0; j < kEpt; ++j)
        hit[j] = eb + j < e1 && !full_l[li[j]] && in_bitmap(a.fb, u[j]);
      // phase 3: lane words of the hits
      uint64_t x[kEpt][K];
#pragma unroll
      for (int j = 0; j < kEpt; ++j) {
        if (hit[j]) {
          load_lanes<K>(a.fl + (int64_t)u[j] * K, x[j]);
        } else {
#pragma unroll
          for (int k = 0; k < K; ++k) x[j][k] = 0;
        }
      }
      // phase 4: OR runs of equal owner, one LDS atomic per run
      uint64_t run[K];
#pragma unroll
      for (int k = 0; k < K; ++k) run[k] = x[0][k];
#pragma unroll
      for (int j = 1; j < kEpt; ++j) {
        if (li[j] != li[j - 1]) stream_flush<K>(acc_l, li[j - 1], run);
#pragma unroll
        for (int k = 0; k < K; ++k) run[k] |= x[j][k];
      }
      stream_flush<K>(acc_l, li[kEpt - 1], run);
    }
    __syncthreads();
    for (int64_t i = t; i < wn; i += kT) {
      const int64_t v = v0 + i;
      const bool inside = win[i] >= e0 && win[i + 1] <= e1;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const unsigned long long x = acc_l[i * K + k];
        if (x) {
          if (inside) a.fn[v * K + k] = x;
          else atomicOr((unsigned long long*)&a.fn[v * K + k], x);
        }
      }
    }
    __syncthreads();
  }
}

// Dense level, hybrid form.  Rows are sorted by target ID, so the targets
// below the hub cut-off `hot` form a prefix of every row: each vertex ORs the
// frontier lanes of that prefix (their lane words are few and stay cached),
// then every frontier vertex at or above the cut-off pushes its lane word
// along its row (mostly into hubs' words, also cached).  Each directed edge
// v <- u is taken once: by v's pull when u < hot, by u's push otherwise.
// F_{L+1} collects the unmasked lanes; k_ms_commit masks them with seen.
template <int K, int G, int U>
__global__ __launch_bounds__(kT) void k_ms_hyb_pull(MsLevel a, int64_t vb, int64_t ve) {
  if (!ms_active(a) || a.dense_algo != kDenseHybrid || ms_mode(a, a.L) != kModeStream) return;
  const int lig = threadIdx.x % G;
  const int64_t gid = ((int64_t)blockIdx.x * kT + threadIdx.x) / G;
  const int64_t ngroups = (int64_t)gridDim.x * (kT / G);
  const uint32_t hot = (uint32_t)a.hot;
  for (int64_t v = vb + gid; v < ve; v += ngroups) {
    uint64_t s[K], f[K];
    load_lanes<K>(a.seen + v * K, s);
    load_lanes<K>(a.fl + v * K, f);
    bool full = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      s[k] |= f[k];
      full = full && ((s[k] | ~a.act_l[k]) == ~0ull);
    }
    if (full) continue;
    const int64_t beg = a.row_off[v], end = a.row_off[v + 1];
    uint64_t acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0;
    for (int64_t base = beg; base < end; base += (int64_t)G * U) {
      uint32_t u[U];
      bool past = false;
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int64_t e = base + (int64_t)j * G + lig;
        u[j] = e < end ? ld_col(a.col + e) : 0xFFFFFFFFu;
        past = past || (e < end && u[j] >= hot);
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if (u[j] < hot && in_bitmap(a.fb, u[j])) {
          uint64_t x[K];
          load_lanes<K>(a.fl + (int64_t)u[j] * K, x);
#pragma unroll
          for (int k = 0; k < K; ++k) acc[k] |= x[k];
        }
      }
      if constexpr (G > 1) {
        int pi = past ? 1 : 0;  // the prefix ends inside this round (group-wide)
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) pi |= __shfl_xor(pi, o, kWave);
        past = pi != 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int o = G / 2; o > 0; o >>= 1) acc[k] |= __shfl_xor(acc[k], o, kWave);
      }
      if (past) break;
    }
    if (lig == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (acc[k]) a.fn[v * K + k] = acc[k];
    }
  }
}

// Heavy rows of the hybrid pull, one wave per work item of `chunk` edges.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_hyb_pull_heavy(MsLevel a,
                                                         const uint64_t* __restrict__ item_start,
                                                         const uint32_t* __restrict__ item_vertex,
                                                         int64_t n_items, int chunk) {
  if (!ms_active(a) || a.dense_algo != kDenseHybrid || ms_mode(a, a.L) != kModeStream) return;
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
  const uint32_t hot = (uint32_t)a.hot;
  for (int64_t it = wid; it < n_items; it += nw) {
    const int64_t v = item_vertex[it];
    const int64_t beg = a.row_off[v] + (int64_t)(it - item_start[v]) * chunk;
    const int64_t end = min(a.row_off[v + 1], beg + (int64_t)chunk);
    if (a.col[beg] >= hot) continue;  // the row's hub prefix ended earlier
    uint64_t s[K], f[K];
    load_lanes<K>(a.seen + v * K, s);
    load_lanes<K>(a.fl + v * K, f);
    bool full = true;
#pragma unroll
    for (int k = 0; k < K; ++k) full = full && ((s[k] | f[k] | ~a.act_l[k]) == ~0ull);
    if (full) continue;
    uint64_t acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0;
    for (int64_t base = beg; base < end; base += kWave) {
      const int64_t e = base + lane;
      const uint32_t u = e < end ? ld_col(a.col + e) : 0xFFFFFFFFu;
      if (u < hot && in_bitmap(a.fb, u)) {
        uint64_t x[K];
        load_lanes<K>(a.fl + (int64_t)u * K, x);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] |= x[k];
      }
      if (__any(e < end && u >= hot)) break;
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc[k] |= __shfl_xor(acc[k], o, kWave);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (acc[k]) atomicOr((unsigned long long*)&a.fn[v * K + k], (unsigned long long)acc[k]);
    }
  }
}


// Push side of the hybrid: frontier vertices in [max(vb, hot), ve), G lanes each.
template <int K, int G>
__global__ __launch_bounds__(kT) void k_ms_hyb_push(MsLevel a, int64_t vb, int64_t ve) {
  if (!ms_active(a) || a.dense_algo != kDenseHybrid || ms_mode(a, a.L) != kModeStream) return;
  const int lig = threadIdx.x % G;
  const int64_t gid = ((int64_t)blockIdx.x * kT + threadIdx.x) / G;
  const int64_t ngroups = (int64_t)gridDim.x * (kT / G);
  for (int64_t v = max(vb, a.hot) + gid; v < ve; v += ngroups) {
    if (!in_bitmap(a.fb, v)) continue;
    uint64_t f[K];
    load_lanes<K>(a.fl + v * K, f);
    const int64_t end = a.row_off[v + 1];
    for (int64_t e0 = a.row_off[v] + lig; e0 < end; e0 += (int64_t)G * kPushIlp) {
      uint32_t w[kPushIlp];
      bool ok[kPushIlp];
      push_targets(a.col, e0, G, end, w, ok);
      hyb_push_n<K>(a, w, ok, f);
    }
  }
}

// Push side for heavy rows at or above the cut-off: items [first_item, n_items).
template <int K>
__global__ __launch_bounds__(kT) void k_ms_hyb_push_heavy(MsLevel a,
                                                         const uint64_t* __restrict__ item_start,
                                                         const uint32_t* __restrict__ item_vertex,
                                                         int64_t first_item, int64_t n_items,
                                                         int chunk) {
  if (!ms_active(a) || a.dense_algo != kDenseHybrid || ms_mode(a, a.L) != kModeStream) return;
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
  for (int64_t it = first_item + wid; it < n_items; it += nw) {
    const int64_t v = item_vertex[it];
    if (!in_bitmap(a.fb, v)) continue;
    uint64_t f[K];
    load_lanes<K>(a.fl + v * K, f);
    const int64_t beg = a.row_off[v] + (int64_t)(it - item_start[v]) * chunk;
    const int64_t end = min(a.row_off[v + 1], beg + (int64_t)chunk);
    for (int64_t e0 = beg + lane; e0 < end; e0 += (int64_t)kWave * kPushIlp) {
      uint32_t w[kPushIlp];
      bool ok[kPushIlp];
      push_targets(a.col, e0, kWave, end, w, ok);
      hyb_push_n<K>(a, w, ok, f);
    }
  }
}

// Split level.  A root whose level-L frontier is light (its degree sum below
// split_thr) would keep almost every vertex's pull scanning its whole row:
// few vertices have a neighbour in that frontier, and a pull stops only
// when every lane it lacks is covered.  Those roots push instead, along the
// rows of their frontier vertices, and the pulls stop once the other lanes
// are covered.
//
// Per-lane degree sums of F_L come from the count pass of the top-down level
// that produced it (k_ms_td_count).
// One block: sums the lane_part rows per lane (each wave a quarter of the
// rows, 16 loads in flight per lane) and sets split_mask to the lanes
// present in F_L whose frontier degree sum is below split_thr.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_lane_mask(MsLevel a, int parts) {
  __shared__ unsigned long long mk[4];
  __shared__ unsigned long long ws[kWavesPerBlock][64 * K];
  if (!ms_active(a) || ms_mode(a, a.L) != kModeSplit) return;
  const int t = threadIdx.x, wave = t / kWave, lane = lane_id();
  if (t < 4) mk[t] = 0;
  const int per = (parts + kWavesPerBlock - 1) / kWavesPerBlock;
  const int p0 = wave * per, p1 = min(parts, p0 + per);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int lanei = 64 * k + lane;
    unsigned long long acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0;
    for (int p = p0; p < p1; p += 16) {
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (p + j < p1) acc[j] += a.lane_part[(int64_t)(p + j) * 256 + lanei];
    }
    unsigned long long sum = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) sum += acc[j];
    ws[wave][lanei] = sum;
  }
  __syncthreads();
  if (t < 64 * K) {
    unsigned long long sum = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) sum += ws[w][t];
    const bool present = (a.act_l[t >> 6] >> (t & 63)) & 1ull;
    if (present && sum < a.split_thr) atomicOr(&mk[t >> 6], 1ull << (t & 63));
  }
  __syncthreads();
  if (t < 4) a.split_mask[t] = mk[t];
}


// After a streaming level: F_{L+1} = gathered lanes & ~seen, seen |= F_{L+1},
// frontier bitmap, counters.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_commit(MsLevel a, int want) {
  __shared__ uint64_t red[kWavesPerBlock];
  if (!ms_active(a) || ms_mode(a, a.L) != want) return;
  uint64_t disc = 0, edges = 0, lanes[K];
#pragma unroll
  for (int k = 0; k < K; ++k) lanes[k] = 0;
  const int64_t stride = (int64_t)gridDim.x * kT;
  for (int64_t vbase = (int64_t)blockIdx.x * kT + (threadIdx.x / kWave) * kWave; vbase < a.n_c;
       vbase += stride) {
    const int64_t v = vbase + lane_id();
    bool nz = false;
    if (v < a.n_c) {
      uint64_t s0[K], f[K], x[K];
      load_lanes<K>(a.seen + v * K, s0);
      load_lanes<K>(a.fl + v * K, f);
      load_lanes<K>(a.fn + v * K, x);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint64_t s = s0[k] | f[k];
        const uint64_t nxt = x[k] & ~s;
        if (nxt != x[k]) a.fn[v * K + k] = nxt;
        if ((s | nxt) != s0[k]) a.seen[v * K + k] = s | nxt;
        disc += (uint64_t)__popcll(nxt);
        lanes[k] |= nxt;
        nz = nz || nxt != 0;
      }
      if (nz) edges += a.deg[v];
    }
    // 64 consecutive vertices from a multiple of 64: one bitmap word
    const uint64_t word = __ballot(nz);
    if (lane_id() == 0 && word) atomicOr((unsigned long long*)&a.fbn[vbase >> 6], (unsigned long long)word);
  }
  ms_note_lanes<K>(a, lanes);
  ms_flush(a, red, disc, edges);
}

// Bottom-up over new IDs [vb, ve), G lanes per vertex, U neighbours per lane
// per round; a vertex stops at the first round after which every lane it
// lacks has a neighbour in that lane's frontier.  The neighbours' lane words
// are read without a frontier-bitmap test first (a word outside the
// frontier is zero): on the levels that pull, most neighbours are in the
// frontier, and the test cost a round trip per round.
template <int K, int G, int U>
__global__ __launch_bounds__(kT) void k_ms_bu(MsLevel a, int64_t vb, int64_t ve) {
  __shared__ uint64_t red[kWavesPerBlock];
  if (!ms_active(a)) return;
  const int mode = ms_mode(a, a.L);
  if (mode != kModePull && mode != kModeSplit) return;
  // split level: the pushing lanes count as covered too, and the lanes the
  // pushes (queued first) put into F_{L+1} join the lanes pulled
  const bool split = mode == kModeSplit;
  const int lig = threadIdx.x % G;
  const int64_t gid0 = ((int64_t)blockIdx.x * kT + threadIdx.x) / G;
  const int64_t ngroups = (int64_t)gridDim.x * (kT / G);
  uint64_t disc = 0, edges = 0;
  // a lane absent from F_L cannot be found this level: it counts as covered
  // when deciding that a vertex is done (roots whose search has ended, e.g.
  // in a small component, would otherwise keep every row scanning to its end)
  uint64_t inact[K], lanes[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    inact[k] = ~a.act_l[k] | (split ? a.split_mask[k] : 0ull);
    lanes[k] = 0;
  }
  // the loop runs wave-uniformly (lanes past ve idle) so that the frontier
  // bitmap bits of a wave's vertices can be combined with shuffles
  const int64_t lane_span = (int64_t)kWave / G;  // vertices per wave
  const int64_t wave_first = gid0 - (lane_id() / G);
  for (int64_t vbase = vb + wave_first; vbase < ve; vbase += ngroups) {
    const int64_t v = vbase + lane_id() / G;
    const bool valid = v < ve;
    uint64_t nxt[K], acc[K];
    bool nz = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      nxt[k] = 0;
      acc[k] = 0;
    }
    if (valid) {
      uint64_t s0[K], f[K], s[K];
      load_lanes<K>(a.seen + v * K, s0);
      load_lanes<K>(a.fl + v * K, f);
      if (split) load_lanes<K>(a.fn + v * K, acc);
      bool full = true;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        s[k] = s0[k] | f[k];
        full = full && ((s[k] | inact[k]) == ~0ull);
      }
      if (!full) {
        const int64_t beg = a.row_off[v], end = a.row_off[v + 1];
        for (int64_t base = beg; base < end; base += (int64_t)G * U) {
          uint32_t u[U];
#pragma unroll
          for (int j = 0; j < U; ++j) {
            const int64_t e = base + (int64_t)j * G + lig;
            u[j] = e < end ? ld_col(a.col + e) : 0xFFFFFFFFu;
          }
#pragma unroll
          for (int j = 0; j < U; ++j) {
            if (u[j] != 0xFFFFFFFFu) {
              uint64_t x[K];
              load_lanes<K>(a.fl + (int64_t)u[j] * K, x);
#pragma unroll
              for (int k = 0; k < K; ++k) acc[k] |= x[k];
            }
          }
          if constexpr (G > 1) {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
              for (int o = G / 2; o > 0; o >>= 1) acc[k] |= __shfl_xor(acc[k], o, kWave);
          }
          bool cov = true;
#pragma unroll
          for (int k = 0; k < K; ++k) cov = cov && ((acc[k] | s[k] | inact[k]) == ~0ull);
          if (cov) break;
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        nxt[k] = acc[k] & ~s[k];
        nz = nz || nxt[k] != 0;
      }
      if (lig == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          a.fn[v * K + k] = nxt[k];
          const uint64_t ns = s[k] | nxt[k];
          if (ns != s0[k]) a.seen[v * K + k] = ns;
          disc += (uint64_t)__popcll(nxt[k]);
          lanes[k] |= nxt[k];
        }
        if (nz) edges += a.deg[v];
      }
    }
    // frontier bitmap of the wave's vertices: they are consecutive, so they
    // touch at most two bitmap words (lane_span <= 64)
    const uint64_t mybit = (valid && nz && lig == 0) ? (1ull << (v & 63)) : 0ull;
    const int64_t w0 = (vbase) >> 6;
    uint64_t lo = ((v >> 6) == w0) ? mybit : 0ull, hi = ((v >> 6) != w0) ? mybit : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo |= __shfl_xor(lo, o, kWave);
      hi |= __shfl_xor(hi, o, kWave);
    }
    if (lane_id() == 0) {
      if (lo) atomicOr((unsigned long long*)&a.fbn[w0], (unsigned long long)lo);
      if (hi) atomicOr((unsigned long long*)&a.fbn[w0 + 1], (unsigned long long)hi);
    }
    (void)lane_span;
  }
  ms_note_lanes<K>(a, lanes);
  ms_flush(a, red, disc, edges);
}

// Bottom-up for heavy vertices: one wave per work item of `chunk` edges ORs
// the frontier lanes it finds into acc[v]; k_ms_bu_heavy_fin turns acc into
// F_{L+1} once every item is done.
template <int K, int U>
__global__ __launch_bounds__(kT) void k_ms_bu_heavy(MsLevel a,
                                                   const uint64_t* __restrict__ item_start,
                                                   const uint32_t* __restrict__ item_vertex,
                                                   int64_t n_items, int chunk,
                                                   uint64_t* __restrict__ gacc) {
  if (!ms_active(a)) return;
  const int mode = ms_mode(a, a.L);
  if (mode != kModePull && mode != kModeSplit) return;
  const int lane = lane_id();
  const int64_t wid = ((int64_t)blockIdx.x * kT + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
  // lanes absent from F_L (and on a split level the pushing lanes) count as
  // covered (as in k_ms_bu)
  uint64_t inact[K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    inact[k] = ~a.act_l[k] | (mode == kModeSplit ? a.split_mask[k] : 0ull);
  for (int64_t it = wid; it < n_items; it += nw) {
    const int64_t v = item_vertex[it];
    uint64_t s[K], g0[K], f[K];
    load_lanes<K>(a.seen + v * K, s);
    load_lanes<K>(a.fl + v * K, f);
    load_lanes<K>(gacc + v * K, g0);
    bool full = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      s[k] |= f[k];
      full = full && ((s[k] | g0[k] | inact[k]) == ~0ull);
    }
    if (full) continue;
    const int64_t beg = a.row_off[v] + (int64_t)(it - item_start[v]) * chunk;
    const int64_t end = min(a.row_off[v + 1], beg + (int64_t)chunk);
    uint64_t acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0;
    for (int64_t base = beg; base < end; base += (int64_t)kWave * U) {
      uint32_t u[U];
#pragma unroll
      for (int j = 0; j < U; ++j) {
        const int64_t e = base + (int64_t)j * kWave + lane;
        u[j] = e < end ? ld_col(a.col + e) : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int j = 0; j < U; ++j) {
        if (u[j] != 0xFFFFFFFFu) {
          uint64_t x[K];
          load_lanes<K>(a.fl + (int64_t)u[j] * K, x);
#pragma unroll
          for (int k = 0; k < K; ++k) acc[k] |= x[k];
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc[k] |= __shfl_xor(acc[k], o, kWave);
      bool cov = true;
#pragma unroll
      for (int k = 0; k < K; ++k) cov = cov && ((acc[k] | s[k] | g0[k] | inact[k]) == ~0ull);
      if (cov) break;
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint64_t nb = acc[k] & ~s[k] & ~g0[k];
        if (nb) atomicOr((unsigned long long*)&gacc[v * K + k], (unsigned long long)nb);
      }
    }
  }
}

template <int K>
__global__ __launch_bounds__(kT) void k_ms_bu_heavy_fin(MsLevel a, int64_t h,
                                                       uint64_t* __restrict__ gacc) {
  __shared__ uint64_t red[kWavesPerBlock];
  if (!ms_active(a)) return;
  const int mode = ms_mode(a, a.L);
  if (mode != kModePull && mode != kModeSplit) return;
  uint64_t disc = 0, edges = 0, lanes[K];
#pragma unroll
  for (int k = 0; k < K; ++k) lanes[k] = 0;
  for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < h;
       v += (int64_t)gridDim.x * kT) {
    uint64_t s0[K], f[K], x[K];
    load_lanes<K>(a.seen + v * K, s0);
    load_lanes<K>(a.fl + v * K, f);
    load_lanes<K>(gacc + v * K, x);
    if (mode == kModeSplit) {  // the lanes the pushes put there join
      uint64_t p[K];
      load_lanes<K>(a.fn + v * K, p);
#pragma unroll
      for (int k = 0; k < K; ++k) x[k] |= p[k];
    }
    bool nz = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (x[k]) gacc[v * K + k] = 0;
      const uint64_t s = s0[k] | f[k];
      const uint64_t nxt = x[k] & ~s;
      a.fn[v * K + k] = nxt;
      if ((s | nxt) != s0[k]) a.seen[v * K + k] = s | nxt;
      disc += (uint64_t)__popcll(nxt);
      lanes[k] |= nxt;
      nz = nz || nxt != 0;
    }
    if (nz) {
      atomicOr((unsigned long long*)&a.fbn[v >> 6], 1ull << (v & 63));
      edges += a.deg[v];
    }
  }
  ms_note_lanes<K>(a, lanes);
  ms_flush(a, red, disc, edges);
}

// Number of levels with a non-empty frontier, or -1 when the traversal may
// continue past the last level buffer (the host completes it).
__device__ inline int ms_levels(const unsigned long long* cnt, int cap) {
  for (int l = 0; l < cap; ++l)
    if (cnt[l] == 0) return l;
  return -1;
}

// bit j of byte i -> bit i of byte j (8x8 bit-matrix transpose)
__device__ inline uint64_t transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

// out[r][o] = level of original vertex o for root r (0xFF unreached).  One
// lane per original ID, so each root's row is written 64 bytes at a time.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_output(const int64_t* __restrict__ d_roots,
                                                 const uint32_t* __restrict__ o2n,
                                                 int64_t n, int64_t n_c,
                                                 const uint64_t* const* __restrict__ fptr,
                                                 const unsigned long long* __restrict__ cnt,
                                                 int cap, uint8_t* __restrict__ out) {
  const int nroots = (int)d_roots[0];
  const int nlev = ms_levels(cnt, cap);
  if (nlev < 0) return;
  for (int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x; o < n;
       o += (int64_t)gridDim.x * kT) {
    const int64_t v = o2n[o];
    uint64_t P[8][K], R[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      R[k] = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) P[b][k] = 0;
    }
    if (v < n_c) {
      for (int l = 0; l < nlev; ++l) {
        uint64_t f[K];
        load_lanes<K>(fptr[l] + v * K, f);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          R[k] |= f[k];
#pragma unroll
          for (int b = 0; b < 8; ++b)
            if ((l >> b) & 1) P[b][k] |= f[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int b = 0; b < 8; ++b) P[b][k] |= ~R[k];
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int sub = 0; sub < 8; ++sub) {
        const int r0 = 64 * k + 8 * sub;
        if (r0 < nroots) {
          uint64_t x = 0;
#pragma unroll
          for (int b = 0; b < 8; ++b) x |= ((P[b][k] >> (8 * sub)) & 0xFFull) << (8 * b);
          const uint64_t y = transpose8(x);
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (r0 + j < nroots) out[(int64_t)(r0 + j) * n + o] = (uint8_t)(y >> (8 * j));
        }
      }
    }
  }
}

// The same through an LDS transpose (n a multiple of kT): a block takes kT
// consecutive original IDs, one per thread, builds the bit planes of all K
// root words from one pass over the vertex's level frontiers, then, one root
// word at a time, writes the 64 levels of each vertex into an LDS tile
// [64 roots][kT vertices] and stores each root's kT-byte row segment with
// 16-byte stores.
//
// With one or two root words and at most kPlaneLevels levels, k_ms_planes
// first walks the vertices in new-ID order (each level frontier read once,
// coalesced) and stores every vertex's levels as four bit planes, contiguous
// per vertex (4 K words: 32 or 64 bytes; 15 = not reached); the output then
// reads that one run per vertex instead of a word per level at a random
// place in each frontier (the level reads fetched ~4x their bytes).
constexpr int kPlaneLevels = 15;

template <int K>
__global__ __launch_bounds__(kT) void k_ms_planes(const uint64_t* const* __restrict__ fptr,
                                                 const unsigned long long* __restrict__ cnt,
                                                 int cap, int64_t n_c,
                                                 uint64_t* __restrict__ planes) {
  const int nlev = ms_levels(cnt, cap);
  if (nlev < 0 || nlev > kPlaneLevels) return;
  for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n_c;
       v += (int64_t)gridDim.x * kT) {
    uint64_t P[4][K], R[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      R[k] = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) P[b][k] = 0;
    }
    for (int l = 0; l < nlev; ++l) {
      uint64_t f[K];
      load_lanes<K>(fptr[l] + v * K, f);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        R[k] |= f[k];
#pragma unroll
        for (int b = 0; b < 4; ++b)
          if ((l >> b) & 1) P[b][k] |= f[k];
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int k = 0; k < K; ++k) planes[v * 4 * K + b * K + k] = P[b][k] | ~R[k];
  }
}

template <int K>
__global__ __launch_bounds__(kT) void k_ms_output_tile(const int64_t* __restrict__ d_roots,
                                                      const uint32_t* __restrict__ o2n,
                                                      int64_t n, int64_t n_c,
                                                      const uint64_t* const* __restrict__ fptr,
                                                      const unsigned long long* __restrict__ cnt,
                                                      int cap, const uint64_t* __restrict__ planes,
                                                      uint8_t* __restrict__ out) {
  __shared__ __align__(16) uint8_t tile[64 * kT];
  const int nroots = (int)d_roots[0];
  const int nlev = ms_levels(cnt, cap);
  if (nlev < 0) return;
  if (nlev > kPlaneLevels) planes = nullptr;  // k_ms_planes left them
  const int t = threadIdx.x;
  for (int64_t base = (int64_t)blockIdx.x * kT; base < n; base += (int64_t)gridDim.x * kT) {
    const int64_t v = o2n[base + t];
    uint64_t P[8][K], R[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      R[k] = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) P[b][k] = 0;
    }
    if (v < n_c && planes) {
      // four planes; 15 (all four set) marks a lane that never reached v,
      // whose upper planes are set too (0xFF)
#pragma unroll
      for (int b = 0; b < 4; ++b) load_lanes<K>(planes + v * 4 * K + b * K, P[b]);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const uint64_t none = P[0][k] & P[1][k] & P[2][k] & P[3][k];
        R[k] = ~none;
      }
    } else if (v < n_c) {
      int l = 0;
      for (; l + 2 <= nlev; l += 2) {
        uint64_t f0[K], f1[K];
        load_lanes<K>(fptr[l] + v * K, f0);
        load_lanes<K>(fptr[l + 1] + v * K, f1);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          R[k] |= f0[k] | f1[k];
#pragma unroll
          for (int b = 0; b < 8; ++b)
            P[b][k] |= (((l >> b) & 1) ? f0[k] : 0ull) | ((((l + 1) >> b) & 1) ? f1[k] : 0ull);
        }
      }
      if (l < nlev) {
        uint64_t f0[K];
        load_lanes<K>(fptr[l] + v * K, f0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          R[k] |= f0[k];
#pragma unroll
          for (int b = 0; b < 8; ++b)
            if ((l >> b) & 1) P[b][k] |= f0[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int b = 0; b < 8; ++b) P[b][k] |= ~R[k];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (64 * k < nroots) {
#pragma unroll
        for (int sub = 0; sub < 8; ++sub) {
          uint64_t x = 0;
#pragma unroll
          for (int b = 0; b < 8; ++b) x |= ((P[b][k] >> (8 * sub)) & 0xFFull) << (8 * b);
          const uint64_t y = transpose8(x);
#pragma unroll
          for (int j = 0; j < 8; ++j) tile[(8 * sub + j) * kT + t] = (uint8_t)(y >> (8 * j));
        }
        __syncthreads();
        for (int c = t; c < 64 * (kT / 16); c += kT) {
          const int row = c / (kT / 16), c16 = c % (kT / 16);
          const int r = 64 * k + row;
          if (r < nroots)
            *reinterpret_cast<uint4*>(out + (int64_t)r * n + base + 16 * c16) =
                *reinterpret_cast<const uint4*>(tile + row * kT + 16 * c16);
        }
        __syncthreads();
      }
    }
  }
}

// levels[i] = level of original vertex pairs[2i+1] for batch root pairs[2i],
// read from the level frontiers: the l with bit r of F_l[v] (0xFF if none or
// if the vertex is outside [0, n)); a root without edges is at level 0.
__global__ void k_ms_gather(const int64_t* __restrict__ d_roots,
                            const uint32_t* __restrict__ o2n, int64_t n, int64_t n_c, int K,
                            const uint64_t* const* __restrict__ fptr,
                            const unsigned long long* __restrict__ cnt, int cap,
                            const int64_t* __restrict__ pairs, int64_t count,
                            uint8_t* __restrict__ levels) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t r = pairs[2 * i], o = pairs[2 * i + 1];
  uint8_t lev = 0xFF;
  if (o >= 0 && o < n) {
    const int64_t v = o2n[o];
    if (v < n_c) {
      const int nlev = ms_levels(cnt, cap);
      const uint64_t bit = 1ull << (r & 63);
      for (int l = 0; l < nlev; ++l)
        if (fptr[l][v * K + (r >> 6)] & bit) {
          lev = (uint8_t)l;
          break;
        }
    } else if (d_roots[1 + r] == o) {
      lev = 0;
    }
  }
  levels[i] = lev;
}

// Roots without edges reach only themselves.
__global__ void k_ms_isolated(const int64_t* __restrict__ d_roots,
                              const uint32_t* __restrict__ o2n, int64_t n, int64_t n_c,
                              const uint32_t* __restrict__ cnt0,
                              const unsigned long long* __restrict__ cnt, int cap,
                              uint8_t* __restrict__ out, unsigned long long* __restrict__ te) {
  const int nroots = (int)d_roots[0];
  const int i = threadIdx.x;
  if (i >= nroots || ms_levels(cnt, cap) < 0) return;
  const int64_t o = d_roots[1 + i];
  const int64_t v = o2n[o];
  if (v < n_c) return;
  if (out) out[(int64_t)i * n + o] = 0;
  te[i] = cnt0[v];
}

// te[r] = sum of cnt0 over the vertices lane r reached.  Words where every
// valid lane (or most lanes) is set take one add per thread; the rest add or
// subtract per bit in LDS.
template <int K>
__global__ __launch_bounds__(kT) void k_ms_te(const int64_t* __restrict__ d_roots,
                                             const uint64_t* __restrict__ seen,
                                             const uint32_t* __restrict__ cnt0, int64_t n_c,
                                             const unsigned long long* __restrict__ cnt,
                                             int cap, unsigned long long* __restrict__ te) {
  __shared__ unsigned long long lane_sum[64 * K];
  __shared__ uint64_t red[kWavesPerBlock];
  const int nroots = (int)d_roots[0];
  if (ms_levels(cnt, cap) < 0) return;
  for (int i = threadIdx.x; i < 64 * K; i += kT) lane_sum[i] = 0;
  __syncthreads();
  uint64_t valid[K], full[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    valid[k] = ~pad_mask(k, nroots);
    full[k] = 0;
  }
  for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n_c;
       v += (int64_t)gridDim.x * kT) {
    const uint64_t c = cnt0[v];
    if (c == 0) continue;
    uint64_t w[K];
    load_lanes<K>(seen + v * K, w);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t x = w[k] & valid[k];
      if (x == valid[k]) {
        full[k] += c;
      } else if (__popcll(x) * 2 > __popcll(valid[k])) {
        full[k] += c;
        for (uint64_t m = valid[k] & ~x; m; m &= m - 1)
          atomicAdd(&lane_sum[64 * k + __ffsll((long long)m) - 1], (unsigned long long)(0 - c));
      } else {
        for (uint64_t m = x; m; m &= m - 1)
          atomicAdd(&lane_sum[64 * k + __ffsll((long long)m) - 1], (unsigned long long)c);
      }
    }
  }
  __shared__ uint64_t full_tot[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t s = block_sum(full[k], red);
    if (threadIdx.x == 0) full_tot[k] = s;
    __syncthreads();
  }
  for (int r = threadIdx.x; r < nroots; r += kT) {
    const unsigned long long tot = lane_sum[r] + full_tot[r >> 6];
    if (tot) atomicAdd(&te[r], tot);
  }
}

int num_cus(int device) {
  int n = 0;
  G500_CHECK_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device));
  return std::max(n, 1);
}

unsigned sweep_grid(const g500_msbfs* b, int64_t groups, int G) {
  const int64_t cap = (int64_t)num_cus(b->g->device) * b->tuning.blocks_per_cu;
  const int64_t need = ceil_div(groups * G, kT);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cap, need));
}

int64_t bitmap_words(const g500_msbfs* b) { return std::max<int64_t>(1, ceil_div(b->g->n_c, 64)); }

MsLevel level_args(const g500_msbfs* b, int L) {
  const g500_graph* g = b->g;
  MsLevel a;
  a.row_off = g->row_off.get();
  a.col = g->col.get();
  a.deg = g->deg_new.get();
  a.seen = b->seen.get();
  a.fl = b->F[L]->get();
  a.fn = b->F[L + 1]->get();
  a.fb = b->fb[L & 1].get();
  a.fbn = b->fb[(L + 1) & 1].get();
  a.cnt = b->cnt.get();
  a.fe = b->fe.get();
  a.act_l = b->act.get() + (int64_t)L * 4;
  a.act_n = b->act.get() + (int64_t)(L + 1) * 4;
  a.thr = b->td_threshold;
  a.d_roots = b->roots.get();
  a.stream_frac = b->tuning.stream_frac;
  a.dense_algo = b->tuning.dense_algo;
  a.hot = 0;  // set per lane-word count by level_kernels
  a.n_c = g->n_c;
  a.L = L;
  a.split_thr = (unsigned long long)((double)b->tuning.split_bp / 10000.0 * (double)g->m);
  a.lane_part = b->lane_part.get();
  a.split_mask = b->split_mask.get();
  return a;
}

template <int K, int G>
void launch_push(const g500_msbfs* b, const MsLevel& a, int64_t vb, int64_t ve, hipStream_t s) {
  if (ve <= vb) return;
  hipLaunchKernelGGL((k_ms_push<K, G>), dim3(sweep_grid(b, ve - vb, G)), dim3(kT), 0, s, a, vb,
                     ve);
  G500_CHECK_LAUNCH();
}

template <int K, int G, int U>
void launch_bu(const g500_msbfs* b, const MsLevel& a, int64_t vb, int64_t ve, hipStream_t s) {
  if (ve <= vb) return;
  hipLaunchKernelGGL((k_ms_bu<K, G, U>), dim3(sweep_grid(b, ve - vb, G)), dim3(kT), 0, s, a, vb,
                     ve);
  G500_CHECK_LAUNCH();
}

// One level L -> L+1: preparation, then the top-down and the bottom-up
// kernels of every bin (the ones of the other direction return at once).
template <int K, int G>
void launch_hyb(const g500_msbfs* b, const MsLevel& a, int64_t vb, int64_t ve, hipStream_t s) {
  if (ve <= vb) return;
  constexpr int U = G == 1 ? 4 : 2;
  hipLaunchKernelGGL((k_ms_hyb_pull<K, G, U>), dim3(sweep_grid(b, ve - vb, G)), dim3(kT), 0, s,
                     a, vb, ve);
  G500_CHECK_LAUNCH();
}

template <int K, int G>
void launch_hyb_push(const g500_msbfs* b, const MsLevel& a, int64_t vb, int64_t ve,
                     hipStream_t s) {
  vb = std::max(vb, a.hot);
  if (ve <= vb) return;
  hipLaunchKernelGGL((k_ms_hyb_push<K, G>), dim3(sweep_grid(b, ve - vb, G)), dim3(kT), 0, s, a,
                     vb, ve);
  G500_CHECK_LAUNCH();
}

template <int K>
void level_kernels(const g500_msbfs* b, int L, hipStream_t s) {
  MsLevel a = level_args(b, L);
  a.hot = std::min<int64_t>(b->g->n_c, b->tuning.hot_bytes / (8 * K));
  // with more than two lane words a pushed lane costs more atomics per
  // edge and a pulled vertex's scan covers more lanes at once: the push
  // threshold shrinks as 2/K (measured best at SCALE 26: 200 basis points
  // for 64 and 128 roots, 50-100 for 256)
  if (K > 2) a.split_thr = (unsigned long long)((double)a.split_thr * 2.0 / K);
  const int64_t h = b->bin_heavy, m = b->bin_medium, n_c = b->g->n_c;
  const int chunk = b->tuning.heavy_chunk;
  const int64_t bw = bitmap_words(b);
  const unsigned pgrid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(n_c * K, kT), (int64_t)num_cus(b->g->device) * 8));
  hipLaunchKernelGGL((k_ms_prep<K>), dim3(pgrid), dim3(kT), 0, s, a, bw);
  G500_CHECK_LAUNCH();
  // a split level's pushing lanes (split levels need split_bp > 0 and come
  // after a top-down level)
  if (a.split_thr > 0 && L > 0) {
    hipLaunchKernelGGL((k_ms_lane_mask<K>), dim3(1), dim3(kT), 0, s, a, kLaneParts);
    G500_CHECK_LAUNCH();
  }
  // pushes: every lane at a top-down level, the pushing lanes at a split
  // level (before the pulls, which take the pushed lanes into F_{L+1})
  const unsigned hgrid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(b->n_items, kWavesPerBlock),
                           (int64_t)num_cus(b->g->device) * b->tuning.blocks_per_cu));
  if (b->n_items > 0) {
    hipLaunchKernelGGL((k_ms_push_heavy<K>), dim3(hgrid), dim3(kT), 0, s, a,
                       b->item_start.get(), b->item_vertex.get(), b->n_items, chunk);
    G500_CHECK_LAUNCH();
  }
  switch (b->tuning.medium_group) {
    case 2: launch_push<K, 2>(b, a, h, m, s); break;
    case 4: launch_push<K, 4>(b, a, h, m, s); break;
    case 8: launch_push<K, 8>(b, a, h, m, s); break;
    case 16: launch_push<K, 16>(b, a, h, m, s); break;
    default: launch_push<K, 32>(b, a, h, m, s); break;
  }
  launch_push<K, 1>(b, a, m, n_c, s);
  hipLaunchKernelGGL((k_ms_td_count<K>), dim3(sweep_grid(b, n_c, 1)), dim3(kT), 0, s, a);
  G500_CHECK_LAUNCH();
  // Every kernel below is queued for every level and returns at once unless
  // the level's mode (decided on the device) is its own; such an empty
  // kernel still costs ~4-5 us.  Streaming levels need stream_frac > 0, and
  // split levels split_bp > 0, so without them their kernels are left out.
  if (b->g->m > 0 && b->tuning.stream_frac > 0) {
    const int64_t ntiles = ceil_div(b->g->m, StreamTile<K>::kTile);
    const unsigned sgrid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(ntiles, (int64_t)num_cus(b->g->device) * b->tuning.blocks_per_cu));
    hipLaunchKernelGGL((k_ms_bu_stream<K>), dim3(sgrid), dim3(kT), 0, s, a, b->g->m,
                       b->stream_tile_v0.get());
    G500_CHECK_LAUNCH();
    if (b->n_items > 0) {
      hipLaunchKernelGGL((k_ms_hyb_pull_heavy<K>), dim3(hgrid), dim3(kT), 0, s, a,
                         b->item_start.get(), b->item_vertex.get(), b->n_items, chunk);
      G500_CHECK_LAUNCH();
    }
    switch (b->tuning.medium_group) {
      case 2: launch_hyb<K, 2>(b, a, h, m, s); break;
      case 4: launch_hyb<K, 4>(b, a, h, m, s); break;
      case 8: launch_hyb<K, 8>(b, a, h, m, s); break;
      case 16: launch_hyb<K, 16>(b, a, h, m, s); break;
      default: launch_hyb<K, 32>(b, a, h, m, s); break;
    }
    launch_hyb<K, 1>(b, a, m, n_c, s);
    // heavy rows at or above the cut-off push by work item
    const int64_t first_cold = a.hot < h ? (int64_t)b->h_item_start[(size_t)a.hot] : b->n_items;
    if (first_cold < b->n_items) {
      hipLaunchKernelGGL((k_ms_hyb_push_heavy<K>), dim3(hgrid), dim3(kT), 0, s, a,
                         b->item_start.get(), b->item_vertex.get(), first_cold, b->n_items,
                         chunk);
      G500_CHECK_LAUNCH();
    }
    switch (b->tuning.medium_group) {
      case 2: launch_hyb_push<K, 2>(b, a, h, m, s); break;
      case 4: launch_hyb_push<K, 4>(b, a, h, m, s); break;
      case 8: launch_hyb_push<K, 8>(b, a, h, m, s); break;
      case 16: launch_hyb_push<K, 16>(b, a, h, m, s); break;
      default: launch_hyb_push<K, 32>(b, a, h, m, s); break;
    }
    launch_hyb_push<K, 1>(b, a, m, n_c, s);
    hipLaunchKernelGGL((k_ms_commit<K>), dim3(sweep_grid(b, n_c, 1)), dim3(kT), 0, s, a,
                       (int)kModeStream);
    G500_CHECK_LAUNCH();
  }
  if (b->n_items > 0) {
    hipLaunchKernelGGL((k_ms_bu_heavy<K, 4>), dim3(hgrid), dim3(kT), 0, s, a,
                       b->item_start.get(), b->item_vertex.get(), b->n_items, chunk,
                       b->acc.get());
    G500_CHECK_LAUNCH();
    hipLaunchKernelGGL((k_ms_bu_heavy_fin<K>), dim3(sweep_grid(b, h, 1)), dim3(kT), 0, s, a, h,
                       b->acc.get());
    G500_CHECK_LAUNCH();
  }
  switch (b->tuning.medium_group) {
    case 2: launch_bu<K, 2, 4>(b, a, h, m, s); break;
    case 4: launch_bu<K, 4, 4>(b, a, h, m, s); break;
    case 8: launch_bu<K, 8, 4>(b, a, h, m, s); break;
    case 16: launch_bu<K, 16, 2>(b, a, h, m, s); break;
    default: launch_bu<K, 32, 2>(b, a, h, m, s); break;
  }
  launch_bu<K, 1, 4>(b, a, m, n_c, s);
}

void enqueue_level(const g500_msbfs* b, int K, int L, hipStream_t s) {
  switch (K) {
    case 1: level_kernels<1>(b, L, s); break;
    case 2: level_kernels<2>(b, L, s); break;
    case 3: level_kernels<3>(b, L, s); break;
    default: level_kernels<4>(b, L, s); break;
  }
}

// cap: the level buffers the batch's levels were queued for (the output
// kernels skip themselves when the last of them is not empty)
template <int K>
void launch_output(const g500_msbfs* b, hipStream_t s, int cap) {
  const g500_graph* g = b->g;
  if (!b->level_arrays) {
    // point lookups read the level frontiers; no per-root arrays
  } else if (g->n % kT == 0) {
    const unsigned grid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(g->n / kT, (int64_t)num_cus(g->device) * 8));
    const uint64_t* planes = nullptr;
    if (K <= 2 && b->planes.get()) {
      planes = b->planes.get();
      const unsigned pgrid = (unsigned)std::max<int64_t>(
          1, std::min<int64_t>(ceil_div(g->n_c, kT), (int64_t)num_cus(g->device) * 16));
      hipLaunchKernelGGL((k_ms_planes<K>), dim3(pgrid), dim3(kT), 0, s,
                         (const uint64_t* const*)b->fptr.get(), b->cnt.get(), cap, g->n_c,
                         b->planes.get());
      G500_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL((k_ms_output_tile<K>), dim3(grid), dim3(kT), 0, s, b->roots.get(),
                       g->orig2new.get(), g->n, g->n_c, (const uint64_t* const*)b->fptr.get(),
                       b->cnt.get(), cap, planes, b->out.get());
  } else {
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(g->n, kT), 65536);
    hipLaunchKernelGGL((k_ms_output<K>), dim3(grid), dim3(kT), 0, s, b->roots.get(),
                       g->orig2new.get(), g->n, g->n_c, (const uint64_t* const*)b->fptr.get(),
                       b->cnt.get(), cap, b->out.get());
  }
  G500_CHECK_LAUNCH();
  const unsigned tgrid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(g->n_c, kT), (int64_t)num_cus(g->device) * 4));
  hipLaunchKernelGGL((k_ms_te<K>), dim3(tgrid), dim3(kT), 0, s, b->roots.get(),
                     b->seen.get(), b->cnt0.get(), g->n_c, b->cnt.get(), cap, b->te.get());
  G500_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ms_isolated, dim3(1), dim3(G500_MSBFS_MAX_ROOTS), 0, s, b->roots.get(),
                     g->orig2new.get(), g->n, g->n_c, b->cnt0.get(), b->cnt.get(), cap,
                     b->level_arrays ? b->out.get() : nullptr, b->te.get());
  G500_CHECK_LAUNCH();
}

void enqueue_output(const g500_msbfs* b, int K, hipStream_t s, int cap) {
  switch (K) {
    case 1: launch_output<1>(b, s, cap); break;
    case 2: launch_output<2>(b, s, cap); break;
    case 3: launch_output<3>(b, s, cap); break;
    default: launch_output<4>(b, s, cap); break;
  }
}

// The whole batch after the roots are on the device: reset, roots, the
// guarded levels filling the first cap level buffers, outputs.
void enqueue_batch(const g500_msbfs* b, int K, int cap, hipStream_t s) {
  const g500_graph* g = b->g;
  const int64_t words = g->n_c * K;
  const unsigned igrid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(ceil_div(words, kT), (int64_t)num_cus(g->device) * 16));
  hipLaunchKernelGGL(k_ms_init, dim3(igrid), dim3(kT), 0, s, b->roots.get(), g->n_c, K,
                     b->seen.get(), b->F[0]->get(), b->fb[0].get(), b->fb[1].get(),
                     bitmap_words(b), b->acc.get(), b->bin_heavy * (int64_t)K, b->cnt.get(),
                     b->fe.get(), kMaxLevels + 2, b->te.get(), b->act.get());
  G500_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ms_roots, dim3(1), dim3(G500_MSBFS_MAX_ROOTS), 0, s, b->roots.get(),
                     g->orig2new.get(), g->deg_new.get(), g->n_c, K, b->seen.get(),
                     b->F[0]->get(), b->fb[0].get(), b->cnt.get(), b->fe.get(), b->act.get());
  G500_CHECK_LAUNCH();
  for (int L = 0; L + 1 < cap; ++L) enqueue_level(b, K, L, s);
  enqueue_output(b, K, s, cap);
}

void add_level_buffer(g500_msbfs* b) {
  auto* f = new DevBuf<uint64_t>;
  try {
    f->alloc((size_t)std::max<int64_t>(1, b->g->n_c) * b->kmax);
  } catch (...) {
    delete f;
    throw;
  }
  b->F.push_back(f);
  const uint64_t* p = f->get();
  G500_CHECK_HIP(hipMemcpy(b->fptr.get() + (b->F.size() - 1), &p, sizeof(p),
                           hipMemcpyHostToDevice));
}

void drop_graphs(g500_msbfs* b) {
  for (int k = 0; k < 5; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    b->exec[k] = nullptr;
    b->exec_levels[k] = 0;
  }
}

__global__ void k_ms_stream_tiles(const int64_t* __restrict__ row_off, int64_t n_c,
                                  uint32_t* __restrict__ tile_v0) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_c) return;
  const int64_t a = row_off[v], b = row_off[v + 1];
  if (b <= a) return;
  const int64_t t0 = (a + kStreamUnit - 1) / kStreamUnit, t1 = (b - 1) / kStreamUnit;
  for (int64_t t = t0; t <= t1; ++t) tile_v0[t] = (uint32_t)v;
  if (v == n_c - 1)  // entries past the last edge point at the last vertex
    for (int64_t t = t1 + 1; t <= t1 + 2; ++t) tile_v0[t] = (uint32_t)v;
}

__global__ void k_ms_item_vertex(const uint64_t* __restrict__ item_start, int64_t h,
                                 uint32_t* __restrict__ item_vertex) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < h)
    for (uint64_t i = item_start[v]; i < item_start[v + 1]; ++i) item_vertex[i] = (uint32_t)v;
}

__global__ void k_ms_item_counts(const int64_t* __restrict__ row_off, int64_t h, int chunk,
                                 uint32_t* __restrict__ items) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < h) items[v] = (uint32_t)ceil_div(row_off[v + 1] - row_off[v], chunk);
}

// Degree bins, heavy work items and the direction threshold for the tuning.
void update_bins(g500_msbfs* b) {
  g500_graph* g = b->g;
  hipStream_t s = b->stream;
  DevBuf<int64_t> d;
  d.alloc(2);
  hipLaunchKernelGGL(k_ms_bins, dim3(1), dim3(1), 0, s, g->deg_new.get(), g->n_c,
                     b->tuning.heavy_degree, b->tuning.medium_degree, d.get());
  G500_CHECK_LAUNCH();
  int64_t hm[2];
  G500_CHECK_HIP(hipMemcpyAsync(hm, d.get(), sizeof(hm), hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
  b->bin_heavy = hm[0];
  b->bin_medium = std::max(hm[0], hm[1]);
  const int64_t h = b->bin_heavy;
  b->item_start.alloc((size_t)h + 1);
  b->acc.alloc((size_t)std::max<int64_t>(1, h) * b->kmax);
  G500_CHECK_HIP(hipMemsetAsync(b->acc.get(), 0, b->acc.n * sizeof(uint64_t), s));
  b->n_items = 0;
  if (h > 0) {
    DevBuf<uint32_t> items;
    items.alloc((size_t)h);
    hipLaunchKernelGGL(k_ms_item_counts, dim3((unsigned)ceil_div(h, kT)), dim3(kT), 0, s,
                       g->row_off.get(), h, b->tuning.heavy_chunk, items.get());
    G500_CHECK_LAUNCH();
    exclusive_scan<uint32_t>(items.get(), b->item_start.get(), h, b->scan_ws, s);
    uint64_t total = 0;
    G500_CHECK_HIP(hipMemcpyAsync(&total, b->item_start.get() + h, sizeof(total),
                                  hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
    b->n_items = (int64_t)total;
    b->h_item_start.resize((size_t)h + 1);
    G500_CHECK_HIP(hipMemcpy(b->h_item_start.data(), b->item_start.get(),
                             ((size_t)h + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
    b->item_vertex.alloc((size_t)std::max<int64_t>(1, b->n_items));
    hipLaunchKernelGGL(k_ms_item_vertex, dim3((unsigned)ceil_div(h, kT)), dim3(kT), 0, s,
                       b->item_start.get(), h, b->item_vertex.get());
    G500_CHECK_LAUNCH();
    G500_CHECK_HIP(hipStreamSynchronize(s));
  }
  b->td_threshold = (unsigned long long)((double)g->m / b->tuning.td_alpha);
}

}  // namespace

// Staging of the point lookups that allocate nothing: a file-scope device
// array that every GPU holds once for the life of the process.  Host code
// never passes `scratch` itself to a kernel or a copy (on the host its name
// is a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr int64_t kNoAllocPairBytes = G500_NOALLOC_MAX_PAIRS * 2 * (int64_t)sizeof(int64_t);
static __device__ __align__(256) unsigned char
    scratch[kNoAllocPairBytes + G500_NOALLOC_MAX_PAIRS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
unsigned char* device_scratch(int device) {
  if (device < 0 || device >= kMaxDevices)
    fail(G500_ERR_ARG, "device %d outside [0, %d)", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(G500_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
           hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

int64_t msbfs_footprint(const g500_graph* g, int max_roots, bool level_arrays) {
  const int64_t kmax = ceil_div(max_roots, 64);
  const int64_t nc = std::max<int64_t>(1, g->n_c);
  const int64_t bw = std::max<int64_t>(1, ceil_div(g->n_c, 64));
  int64_t bytes = nc * kmax * 8;                      // seen
  bytes += (int64_t)kInitialLevels * nc * kmax * 8;   // level frontiers
  if (level_arrays) {
    bytes += (int64_t)max_roots * g->n;                 // level arrays
    bytes += nc * 4 * std::min<int64_t>(kmax, 2) * 8;  // their bit planes (K <= 2)
  }
  bytes += g->n * 4;                                  // cnt0
  bytes += 2 * bw * 8;                                // frontier bitmaps
  bytes += 4 * (kMaxLevels + 2) * 8 + 8 * 1024;       // counters, roots, te, tables
  bytes += (ceil_div(g->m, kStreamUnit) + 3) * 4;     // streaming tile table
  bytes += (int64_t)kLaneParts * 256 * 8 + 4 * 8;     // split level: lane sums, mask
  return bytes;
}

void msbfs_init(g500_msbfs* b, int max_roots, bool level_arrays) {
  g500_graph* g = b->g;
  b->max_roots = max_roots;
  b->kmax = (int)ceil_div(max_roots, 64);
  (void)device_scratch(g->device);  // the allocation-free lookups' staging
  G500_CHECK_HIP(hipStreamCreate(&b->stream));
  G500_CHECK_HIP(hipEventCreate(&b->ev_begin));
  G500_CHECK_HIP(hipEventCreate(&b->ev_end));
  const int64_t nc = std::max<int64_t>(1, g->n_c);
  b->seen.alloc((size_t)nc * b->kmax);
  b->fptr.alloc(kMaxLevels + 1);  // F_0..F_255
  b->cnt.alloc(kMaxLevels + 2);
  b->te.alloc(G500_MSBFS_MAX_ROOTS);
  b->lane_part.alloc((size_t)kLaneParts * 256);
  b->split_mask.alloc(4);
  G500_CHECK_HIP(hipMemset(b->split_mask.get(), 0, 4 * sizeof(uint64_t)));
  b->roots.alloc(G500_MSBFS_MAX_ROOTS + 1);
  b->level_arrays = level_arrays;
  if (level_arrays) {
    b->out.alloc((size_t)max_roots * (size_t)g->n);
    b->planes.alloc((size_t)nc * 4 * std::min(b->kmax, 2));
  }
  b->fb[0].alloc((size_t)bitmap_words(b));
  b->fb[1].alloc((size_t)bitmap_words(b));
  b->fe.alloc(kMaxLevels + 2);
  b->act.alloc((size_t)(kMaxLevels + 2) * 4);
  b->cnt0.alloc((size_t)g->n);
  G500_CHECK_HIP(hipHostMalloc((void**)&b->h_roots, (G500_MSBFS_MAX_ROOTS + 1) * sizeof(int64_t),
                               hipHostMallocDefault));
  G500_CHECK_HIP(hipHostMalloc((void**)&b->h_cnt, (kMaxLevels + 2) * sizeof(unsigned long long),
                               hipHostMallocDefault));
  for (int l = 0; l < kInitialLevels; ++l) add_level_buffer(b);

  G500_CHECK_HIP(hipMemsetAsync(b->cnt0.get(), 0, (size_t)g->n * sizeof(uint32_t), b->stream));
  if (g->nedges > 0) {
    const unsigned grid = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(ceil_div(g->nedges, kT), (int64_t)num_cus(g->device) * 16));
    hipLaunchKernelGGL(k_ms_cnt0, dim3(grid), dim3(kT), 0, b->stream, g->edges.get(),
                       g->nedges, g->orig2new.get(), b->cnt0.get());
    G500_CHECK_LAUNCH();
  }
  if (g->m > 0) {
    // tiles of 2048 edges read entries 2t and 2t+2: room for both past the end
    b->stream_tile_v0.alloc((size_t)ceil_div(g->m, kStreamUnit) + 3);
    hipLaunchKernelGGL(k_ms_stream_tiles, dim3((unsigned)ceil_div(g->n_c, kT)), dim3(kT), 0,
                       b->stream, g->row_off.get(), g->n_c, b->stream_tile_v0.get());
    G500_CHECK_LAUNCH();
  }
  // defaults: what the autotuner picked at SCALE 26 for 64 and 128 roots
  // (and near what it picks for 256)
  g500_msbfs_tuning t{};
  t.heavy_degree = 8192;
  t.medium_degree = 256;
  t.medium_group = 16;
  t.blocks_per_cu = 16;
  t.use_graph = 1;
  t.heavy_chunk = 8192;
  t.td_alpha = 4.0;
  t.stream_frac = 0.0;  // dense levels off until measured faster (tests force them)
  t.hot_bytes = 2 << 20;
  t.dense_algo = kDenseHybrid;
  t.split_bp = 200;
  msbfs_set_tuning(b, t);
  G500_CHECK_HIP(hipStreamSynchronize(b->stream));
}

void msbfs_set_tuning(g500_msbfs* b, const g500_msbfs_tuning& t) {
  if (t.heavy_degree < 1 || t.medium_degree < 1 || t.medium_degree > t.heavy_degree)
    fail(G500_ERR_ARG, "tuning: need 1 <= medium_degree <= heavy_degree");
  if (t.medium_group != 2 && t.medium_group != 4 && t.medium_group != 8 &&
      t.medium_group != 16 && t.medium_group != 32)
    fail(G500_ERR_ARG, "tuning: medium_group must be 2, 4, 8, 16 or 32");
  if (t.blocks_per_cu < 1 || t.blocks_per_cu > 32)
    fail(G500_ERR_ARG, "tuning: blocks_per_cu must be in [1, 32]");
  if (t.use_graph != 0 && t.use_graph != 1) fail(G500_ERR_ARG, "tuning: use_graph is 0 or 1");
  if (t.heavy_chunk < 256 || t.heavy_chunk > 65536 || t.heavy_chunk % 256)
    fail(G500_ERR_ARG, "tuning: heavy_chunk must be a multiple of 256 in [256, 65536]");
  if (!(t.td_alpha > 0)) fail(G500_ERR_ARG, "tuning: td_alpha must be > 0");
  if (!(t.stream_frac >= 0 && t.stream_frac <= 1))
    fail(G500_ERR_ARG, "tuning: stream_frac must be in [0, 1]");
  if (t.dense_algo != kDenseStream && t.dense_algo != kDenseHybrid)
    fail(G500_ERR_ARG, "tuning: dense_algo is 1 (stream) or 2 (hub pull + push)");
  if (t.hot_bytes < 0) fail(G500_ERR_ARG, "tuning: hot_bytes must be >= 0");
  if (t.split_bp < 0 || t.split_bp > 10000)
    fail(G500_ERR_ARG, "tuning: split_bp must be in [0, 10000]");
  b->tuning = t;
  drop_graphs(b);
  update_bins(b);
}

void msbfs_launch(g500_msbfs* b, const int64_t* roots, int nroots) {
  g500_graph* g = b->g;
  if (b->in_flight) fail(G500_ERR_STATE, "a batch is already in flight (call g500_msbfs_wait)");
  if (!roots || nroots < 1 || nroots > b->max_roots)
    fail(G500_ERR_ARG, "nroots %d outside [1, %d]", nroots, b->max_roots);
  for (int i = 0; i < nroots; ++i)
    if (roots[i] < 0 || roots[i] >= g->n)
      fail(G500_ERR_ARG, "root %lld outside [0, %lld)", (long long)roots[i], (long long)g->n);
  const int K = (int)ceil_div(nroots, 64);
  b->h_roots[0] = nroots;
  std::memcpy(b->h_roots + 1, roots, (size_t)nroots * sizeof(int64_t));
  hipStream_t s = b->stream;
  // levels queued: two more than the deepest batch so far needed (every
  // guarded level step costs its empty launches), all buffers before that
  const int ncap = b->depth_hint > 0 ? std::min(std::max(b->depth_hint, 3), b->levels_cap())
                                     : b->levels_cap();
  if (b->tuning.use_graph && (!b->exec[K] || b->exec_levels[K] != ncap)) {
    // captured once per K and level count; the stream is idle here (no batch
    // in flight), and the capture itself runs nothing
    {
      if (b->exec[K]) (void)hipGraphExecDestroy(b->exec[K]);
      b->exec[K] = nullptr;
      hipGraph_t graph = nullptr;
      G500_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_batch(b, K, ncap, s);
      } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      G500_CHECK_HIP(hipStreamEndCapture(s, &graph));
      const hipError_t e = hipGraphInstantiate(&b->exec[K], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) fail(G500_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      b->exec_levels[K] = ncap;
    }
  }
  G500_CHECK_HIP(hipEventRecord(b->ev_begin, s));
  G500_CHECK_HIP(hipMemcpyAsync(b->roots.get(), b->h_roots, (size_t)(nroots + 1) * sizeof(int64_t),
                                hipMemcpyHostToDevice, s));
  if (b->tuning.use_graph) {
    G500_CHECK_HIP(hipGraphLaunch(b->exec[K], s));
  } else {
    enqueue_batch(b, K, ncap, s);
  }
  b->run_levels = ncap;
  G500_CHECK_HIP(hipEventRecord(b->ev_end, s));
  b->nroots = nroots;
  b->K = K;
  b->in_flight = true;
  b->have_result = false;
}

bool msbfs_query(g500_msbfs* b) {
  if (!b->in_flight) return true;
  const hipError_t e = hipEventQuery(b->ev_end);
  if (e == hipSuccess) return true;
  if (e == hipErrorNotReady) return false;
  fail(G500_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(e));
}

void msbfs_wait(g500_msbfs* b) {
  if (!b->in_flight) {
    if (!b->have_result) fail(G500_ERR_STATE, "no batch launched");
    return;
  }
  hipStream_t s = b->stream;
  b->in_flight = false;
  G500_CHECK_HIP(hipEventSynchronize(b->ev_end));
  float ms = 0;
  G500_CHECK_HIP(hipEventElapsedTime(&ms, b->ev_begin, b->ev_end));
  double gpu_ms = ms;
  const size_t cbytes = (kMaxLevels + 2) * sizeof(unsigned long long);
  G500_CHECK_HIP(hipMemcpyAsync(b->h_cnt, b->cnt.get(), cbytes, hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
  int cap = b->run_levels;
  if (b->h_cnt[cap - 1] != 0) {
    // deeper than the level buffers: continue level by level, then redo the
    // outputs (the captured output kernels skipped themselves)
    G500_CHECK_HIP(hipEventRecord(b->ev_begin, s));
    for (int L = cap - 1;; ++L) {
      while (b->levels_cap() < L + 2) add_level_buffer(b);
      enqueue_level(b, b->K, L, s);
      G500_CHECK_HIP(hipMemcpyAsync(b->h_cnt + L + 1, b->cnt.get() + L + 1,
                                    sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
      G500_CHECK_HIP(hipStreamSynchronize(s));
      if (b->h_cnt[L + 1] == 0) break;
      if (L + 1 >= kMaxLevels)
        fail(G500_ERR_STATE, "a root reaches depth %d: uint8 levels hold depths <= %d", L + 1,
             kMaxLevels - 1);
    }
    // the last buffer now holds the empty frontier, so cnt[cap - 1] == 0
    enqueue_output(b, b->K, s, b->levels_cap());
    G500_CHECK_HIP(hipEventRecord(b->ev_end, s));
    G500_CHECK_HIP(hipEventSynchronize(b->ev_end));
    G500_CHECK_HIP(hipEventElapsedTime(&ms, b->ev_begin, b->ev_end));
    gpu_ms += ms;
    G500_CHECK_HIP(hipMemcpyAsync(b->h_cnt, b->cnt.get(), cbytes, hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
    drop_graphs(b);  // recaptured with the larger level count at next launch
    cap = b->levels_cap();
  }
  g500_msbfs_stats& st = b->stats;
  std::memset(&st, 0, sizeof(st));
  st.nroots = b->nroots;
  st.gpu_ms = gpu_ms;
  int nlev = cap;
  for (int l = 0; l < cap; ++l)
    if (b->h_cnt[l] == 0) {
      nlev = l;
      break;
    }
  st.nlevels = nlev;
  b->depth_hint = std::max(b->depth_hint, nlev + 2);
  for (int l = 0; l < nlev; ++l) {
    st.discovered += (int64_t)b->h_cnt[l];
    if (l < 64) st.lane_discoveries[l] = (int64_t)b->h_cnt[l];
  }
  b->have_result = true;
}

namespace {

void check_pairs(const g500_msbfs* b, const int64_t* pairs, int64_t count) {
  for (int64_t i = 0; i < count; ++i)
    if (pairs[2 * i] < 0 || pairs[2 * i] >= b->nroots)
      fail(G500_ERR_ARG, "pair %lld: root index %lld outside [0, %d)", (long long)i,
           (long long)pairs[2 * i], b->nroots);
}

// Pairs to d_pairs, one thread per pair, levels back from d_levels.
void gather_staged(g500_msbfs* b, const int64_t* pairs, int64_t count, uint8_t* levels,
                   int64_t* d_pairs, uint8_t* d_levels) {
  const g500_graph* g = b->g;
  hipStream_t s = b->stream;
  G500_CHECK_HIP(hipMemcpyAsync(d_pairs, pairs, (size_t)count * 2 * sizeof(int64_t),
                                hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_ms_gather, dim3((unsigned)ceil_div(count, kT)), dim3(kT), 0, s,
                     b->roots.get(), g->orig2new.get(), g->n, g->n_c, b->K,
                     (const uint64_t* const*)b->fptr.get(), b->cnt.get(), b->levels_cap(),
                     d_pairs, count, d_levels);
  G500_CHECK_LAUNCH();
  G500_CHECK_HIP(hipMemcpyAsync(levels, d_levels, (size_t)count, hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
}

}  // namespace

void msbfs_gather_noalloc(g500_msbfs* b, const int64_t* pairs, int64_t count,
                          uint8_t* levels) {
  if (count > G500_NOALLOC_MAX_PAIRS)
    fail(G500_ERR_ARG, "%lld pairs: the allocation-free lookup takes at most %d per call",
         (long long)count, G500_NOALLOC_MAX_PAIRS);
  unsigned char* base = device_scratch(b->g->device);
  if (count <= 0) return;
  check_pairs(b, pairs, count);
  // the kernel writes every one of the `count` level bytes it hands back,
  // so nothing in the scratch needs clearing between calls
  gather_staged(b, pairs, count, levels, reinterpret_cast<int64_t*>(base),
                base + kNoAllocPairBytes);
}

void msbfs_gather(g500_msbfs* b, const int64_t* pairs, int64_t count, uint8_t* levels) {
  if (count <= 0) return;
  check_pairs(b, pairs, count);
  if ((int64_t)b->gather_pairs.n < 2 * count) {
    b->gather_pairs.alloc((size_t)2 * count);
    b->gather_levels.alloc((size_t)count);
  }
  gather_staged(b, pairs, count, levels, b->gather_pairs.get(), b->gather_levels.get());
}

// Coordinate search over the work split and the direction rule: each
// parameter in turn takes every candidate value with the others fixed, the
// fastest batch (device time, best of two runs) is kept; `rounds` passes.
double msbfs_autotune(g500_msbfs* b, const int64_t* roots, int nroots, int rounds,
                      g500_msbfs_tuning* best_out) {
  g500_msbfs_tuning best = b->tuning;
  auto measure = [&](const g500_msbfs_tuning& t) {
    msbfs_set_tuning(b, t);
    double ms = 1e300;
    for (int rep = 0; rep < 3; ++rep) {  // the first run also captures the graph
      msbfs_launch(b, roots, nroots);
      msbfs_wait(b);
      if (rep > 0) ms = std::min(ms, b->stats.gpu_ms);
    }
    return ms;
  };
  double best_ms = measure(best);
  const int64_t heavy[] = {512, 1024, 2048, 4096, 8192, 16384};
  const int64_t medium[] = {8, 16, 32, 64, 128, 256};
  const int32_t group[] = {4, 8, 16, 32};
  const int32_t bpc[] = {2, 4, 8, 16};
  const int32_t chunk[] = {1024, 2048, 4096, 8192, 16384};
  const double alpha[] = {1, 2, 4, 8, 16, 32, 64};
  const double sfrac[] = {0.0, 0.05, 0.2};
  const int64_t hotb[] = {1 << 20, 2 << 20, 4 << 20, 8 << 20, 32 << 20};
  const int32_t dalgo[] = {kDenseStream, kDenseHybrid};
  const int32_t split[] = {0, 50, 100, 200, 400, 800};
  for (int round = 0; round < rounds; ++round) {
    for (int p = 0; p < 10; ++p) {
      static const int ncand[10] = {6, 6, 4, 4, 5, 7, 3, 5, 2, 6};
      const int nc = ncand[p];
      for (int c = 0; c < nc; ++c) {
        g500_msbfs_tuning t = best;
        switch (p) {
          case 0: t.heavy_degree = heavy[c]; break;
          case 1: t.medium_degree = medium[c]; break;
          case 2: t.medium_group = group[c]; break;
          case 3: t.blocks_per_cu = bpc[c]; break;
          case 4: t.heavy_chunk = chunk[c]; break;
          case 5: t.td_alpha = alpha[c]; break;
          case 6: t.stream_frac = sfrac[c]; break;
          case 7: t.hot_bytes = hotb[c]; break;
          case 8: t.dense_algo = dalgo[c]; break;
          default: t.split_bp = split[c]; break;
        }
        if (t.medium_degree > t.heavy_degree) continue;
        if (std::memcmp(&t, &best, sizeof(t)) == 0) continue;
        const double ms = measure(t);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
    }
  }
  msbfs_set_tuning(b, best);
  if (best_out) *best_out = best;
  return best_ms;
}

}  // namespace g500

g500_msbfs::~g500_msbfs() {
  g500::drop_graphs(this);
  for (auto* f : F) delete f;
  if (h_roots) (void)hipHostFree(h_roots);
  if (h_cnt) (void)hipHostFree(h_cnt);
  if (ev_begin) (void)hipEventDestroy(ev_begin);
  if (ev_end) (void)hipEventDestroy(ev_end);
  if (stream) (void)hipStreamDestroy(stream);
}
