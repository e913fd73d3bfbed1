// validate.hip -- the Graph500 result validator on the GPU.
//
// Same checks, same order and the same early exits as
// BfsValidation::validate (mpi/validate.hpp:608-976): parent range, root is
// its own parent, no other self-parent, depth map (or, with CoreBFS, the
// depth rebuild of validate.hpp:1040-1141), then one pass over every input
// edge (tree/non-tree mix, depth gap, parent-edge depth, traversed-edge
// count) and finally "every reached vertex has an edge to its parent".
// One difference: the reference stops its edge pass at the first 8 Mi-edge
// chunk that has errors; this pass always covers all edges, so error totals
// (never the verdict or the count of a valid tree) can differ.
#include "engine.hpp"

namespace g500 {

namespace {

constexpr int kT = 256;
enum { kTE = G500_VERR_NUM, kChanged = G500_VERR_NUM + 1, kCounters = G500_VERR_NUM + 2 };

__device__ inline int64_t pe_parent(int64_t e) {
  return (int64_t)((uint64_t)e << 16) >> 16;
}
__device__ inline uint32_t pe_depth(int64_t e) {
  return (uint32_t)(((uint64_t)e >> 48) & 0xFFFF);
}

__device__ inline uint64_t wsum64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Adds v to cnt[k] with one atomic per wave.
__device__ inline void wave_count(unsigned long long* cnt, int k, uint64_t v) {
  v = wsum64(v);
  if (lane_id() == 0 && v) atomicAdd(&cnt[k], (unsigned long long)v);
}

__global__ __launch_bounds__(kT) void k_v_ranges(const int64_t* __restrict__ pred,
                                                int64_t nlocal, int64_t nglobal,
                                                int64_t root,
                                                unsigned long long* cnt) {
  uint64_t range = 0, self = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < nlocal;
       i += (int64_t)gridDim.x * kT) {
    const int64_t p = pe_parent(pred[i]);
    if (p < -1 || p >= nglobal) ++range;
    if (i != root && p != -1 && p == i) ++self;
  }
  wave_count(cnt, G500_VERR_RANGE, range);
  wave_count(cnt, G500_VERR_SELF_PARENT, self);
}

// validate.hpp:1145-1255 (no CoreBFS)
__global__ __launch_bounds__(kT) void k_v_depths(const int64_t* __restrict__ pred,
                                                int64_t nlocal, int64_t root,
                                                unsigned long long* cnt) {
  uint64_t bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < nlocal;
       i += (int64_t)gridDim.x * kT) {
    const int64_t e = pred[i];
    const int64_t p = pe_parent(e);
    const uint32_t d = pe_depth(e);
    if (p == -1 && d != 0xFFFF) ++bad;
    else if (p != -1 && d == 0xFFFF) ++bad;
    if (i == root || p == -1) continue;
    const uint32_t dp = pe_depth(pred[p]);  // p in range (checked before)
    if (dp == 0xFFFF) ++bad;
    if (d != dp + 1) ++bad;
  }
  wave_count(cnt, G500_VERR_DEPTH_MAP, bad);
}

// validate.hpp:1040-1141: depth[v] = depth[parent]+1 where unknown; one sweep.
__global__ __launch_bounds__(kT) void k_v_depth_build(int64_t* __restrict__ pred,
                                                     int64_t nlocal, int64_t root,
                                                     unsigned long long* cnt) {
  uint64_t changed = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < nlocal;
       i += (int64_t)gridDim.x * kT) {
    if (i == root) continue;
    const int64_t e = pred[i];
    const int64_t p = pe_parent(e);
    if (p < 0 || pe_depth(e) != 0xFFFF) continue;
    const uint32_t dp = pe_depth(__atomic_load_n(&pred[p], __ATOMIC_RELAXED));
    if (dp == 0xFFFF) continue;
    pred[i] = (int64_t)(((uint64_t)e & 0xFFFFFFFFFFFFull) | ((uint64_t)((dp + 1) & 0xFFFF) << 48));
    ++changed;
  }
  wave_count(cnt, kChanged, changed);
}

__global__ void k_v_set_root_depth0(int64_t* pred, int64_t root) {
  pred[root] = (int64_t)((uint64_t)pred[root] & 0xFFFFFFFFFFFFull);
}

__device__ inline int64_t pred_at(const int64_t* pred, int64_t nlocal, int64_t v) {
  return (v >= 0 && v < nlocal) ? pred[v] : (int64_t)-1;
}

// validate.hpp:834-916
__global__ __launch_bounds__(kT) void k_v_edges(const uint2* __restrict__ edges,
                                               int64_t ne,
                                               const int64_t* __restrict__ pred,
                                               int64_t nlocal,
                                               uint8_t* __restrict__ pred_valid,
                                               unsigned long long* cnt) {
  uint64_t te = 0, tree = 0, gap = 0, pdep = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < ne;
       i += (int64_t)gridDim.x * kT) {
    const uint2 x = edges[i];
    const int64_t s = x.x, t = x.y;
    const int64_t ps = pred_at(pred, nlocal, s), pt = pred_at(pred, nlocal, t);
    const int ds = (int)pe_depth(ps), dt = (int)pe_depth(pt);
    if (ds != 0xFFFF && dt == 0xFFFF) ++tree;
    else if (ds == 0xFFFF && dt != 0xFFFF) ++tree;
    else if (ds - dt < -1 || ds - dt > 1) ++gap;
    else if (ds != 0xFFFF) ++te;
    const bool s_to_t = pe_parent(ps) == t, t_to_s = pe_parent(pt) == s;
    if ((s_to_t || t_to_s) && s != t &&
        ((s_to_t && dt + 1 != ds) || (t_to_s && ds + 1 != dt)))
      pdep += (uint64_t)s_to_t + (uint64_t)t_to_s;
    if (s_to_t) pred_valid[s] = 1;
    if (t_to_s) pred_valid[t] = 1;
  }
  wave_count(cnt, kTE, te);
  wave_count(cnt, G500_VERR_EDGE_TREE, tree);
  wave_count(cnt, G500_VERR_EDGE_GAP, gap);
  wave_count(cnt, G500_VERR_PRED_DEPTH, pdep);
}

__global__ __launch_bounds__(kT) void k_v_parent_edges(const int64_t* __restrict__ pred,
                                                      int64_t nlocal, int64_t root,
                                                      const uint8_t* __restrict__ pv,
                                                      unsigned long long* cnt) {
  uint64_t bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < nlocal;
       i += (int64_t)gridDim.x * kT) {
    if (pe_parent(pred[i]) == -1) continue;
    if (!pv[i] && i != root) ++bad;
  }
  wave_count(cnt, G500_VERR_NO_PARENT_EDGE, bad);
}

unsigned grid(int64_t n) {
  const int64_t b = ceil_div(n, kT);
  return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

void bfs_validate(g500_bfs* b, int64_t root, const int64_t* host_pred,
                  int64_t* te, int64_t* errs, int* ok) {
  g500_graph* g = b->g;
  hipStream_t s = b->stream;
  const int64_t nlocal = g->n;
  const int64_t nglobal = g->max_used + 1;
  if (host_pred) {
    G500_CHECK_HIP(hipMemcpyAsync(b->pred.get(), host_pred, nlocal * sizeof(int64_t),
                                  hipMemcpyHostToDevice, s));
    b->pred_fill_needed = true;  // the next BFS run restores the -1 entries
  } else if (!b->have_result) {
    fail(G500_ERR_STATE, "validate: no BFS result on the device");
  }
  if (!b->vcount.get()) b->vcount.alloc(kCounters);
  if (!b->pred_valid.get()) b->pred_valid.alloc((size_t)nlocal);
  unsigned long long* cnt = b->vcount.get();
  unsigned long long h[kCounters];
  auto fetch = [&]() {
    G500_CHECK_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
  };
  for (int i = 0; i < G500_VERR_NUM; ++i) errs[i] = 0;
  *te = 0;
  *ok = 0;
  G500_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(h), s));

  // stage 1-3: ranges, root, self parents
  hipLaunchKernelGGL(k_v_ranges, dim3(grid(nlocal)), dim3(kT), 0, s, b->pred.get(),
                     nlocal, nglobal, root, cnt);
  G500_CHECK_LAUNCH();
  fetch();
  if (h[G500_VERR_RANGE]) {
    errs[G500_VERR_RANGE] = (int64_t)h[G500_VERR_RANGE];
    return;
  }
  if (root < 0 || root >= nlocal) {
    errs[G500_VERR_ROOT] = 1;
    return;
  }
  int64_t proot = 0;
  G500_CHECK_HIP(hipMemcpyAsync(&proot, b->pred.get() + root, sizeof(int64_t),
                                hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t proot_parent = (int64_t)((uint64_t)proot << 16) >> 16;
  if (proot_parent != root) {
    errs[G500_VERR_ROOT] = 1;
    return;
  }
  if (h[G500_VERR_SELF_PARENT]) {
    errs[G500_VERR_SELF_PARENT] = (int64_t)h[G500_VERR_SELF_PARENT];
    return;
  }

  // stage 4: depth map
  if (g->corebfs) {
    hipLaunchKernelGGL(k_v_set_root_depth0, dim3(1), dim3(1), 0, s, b->pred.get(), root);
    G500_CHECK_LAUNCH();
    for (int iter = 0; iter < 0x10000; ++iter) {
      G500_CHECK_HIP(hipMemsetAsync(cnt + kChanged, 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_v_depth_build, dim3(grid(nlocal)), dim3(kT), 0, s,
                         b->pred.get(), nlocal, root, cnt);
      G500_CHECK_LAUNCH();
      fetch();
      if (!h[kChanged]) break;
    }
  } else {
    const uint32_t droot = (uint32_t)(((uint64_t)proot >> 48) & 0xFFFF);
    hipLaunchKernelGGL(k_v_depths, dim3(grid(nlocal)), dim3(kT), 0, s, b->pred.get(),
                       nlocal, root, cnt);
    G500_CHECK_LAUNCH();
    fetch();
    if (droot != 0 || h[G500_VERR_DEPTH_MAP]) {
      errs[G500_VERR_ROOT] = droot != 0;
      errs[G500_VERR_DEPTH_MAP] = (int64_t)h[G500_VERR_DEPTH_MAP];
      return;
    }
  }

  // stage 5: every input edge, then parent edges
  G500_CHECK_HIP(hipMemsetAsync(b->pred_valid.get(), 0, nlocal, s));
  if (g->nedges > 0) {
    hipLaunchKernelGGL(k_v_edges, dim3(grid(g->nedges)), dim3(kT), 0, s,
                       g->edges.get(), g->nedges, b->pred.get(), nlocal,
                       b->pred_valid.get(), cnt);
    G500_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_v_parent_edges, dim3(grid(nlocal)), dim3(kT), 0, s,
                     b->pred.get(), nlocal, root, b->pred_valid.get(), cnt);
  G500_CHECK_LAUNCH();
  fetch();
  for (int i = 0; i < G500_VERR_NUM; ++i) errs[i] = (int64_t)h[i];
  *te = (int64_t)h[kTE];
  int good = 1;
  for (int i = 0; i < G500_VERR_NUM; ++i)
    if (errs[i]) good = 0;
  *ok = good;
}

}  // namespace g500
