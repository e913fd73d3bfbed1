// corebfs.hip -- CoreBFS forest pruning on the GPU.
//
// Preprocess (corebfs_adaptor::preprocess, mpi/corebfs_adaptor.hpp:328-357 ->
// prune_trees, external/indexed_bfs/bfs/corebfs/decomposition.hpp:350-572):
// on the deduplicated symmetric CSR, repeatedly take the vertices of live
// degree 1 ("leaves"); each leaf's parent is its one neighbour that was live
// when the round started; leaves then drop to degree 0 and each parent loses
// one degree, dying when it reaches 0.  What stays live is the 2-core; its
// edges are rebuilt into the CSR the BFS runs on (same vertex space).
//
// reset_unreachable (corebfs_adaptor.hpp:228-249, postprocess.hpp:106-154):
// after one BFS over the core from its highest-degree vertex, a forest vertex
// keeps its parent only if following parents leads into that component.
#include <algorithm>

#include "internal.hpp"

namespace g500 {

namespace {

constexpr int kT = 256;

__device__ inline uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__global__ __launch_bounds__(kT) void k_core_init(const int64_t* __restrict__ row_off,
                                                 int64_t n_c, uint32_t* __restrict__ deg,
                                                 uint8_t* __restrict__ alive,
                                                 uint32_t* __restrict__ par) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n_c) return;
  const uint32_t d = (uint32_t)(row_off[v + 1] - row_off[v]);
  deg[v] = d;
  alive[v] = d > 0;
  par[v] = kNoParent;
}

// leaves (live degree 1) appended to `leaves`; cnt[0] = count
__global__ __launch_bounds__(kT) void k_core_leaves(const uint32_t* __restrict__ deg,
                                                   int64_t n_c,
                                                   uint32_t* __restrict__ leaves,
                                                   unsigned long long* cnt) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  const bool leaf = v < n_c && deg[v] == 1;
  const uint64_t mask = __ballot(leaf);
  if (!mask) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(cnt, (unsigned long long)__popcll(mask));
  base = __shfl(base, 0, kWave);
  if (leaf) leaves[base + __popcll(mask & lanemask_lt())] = (uint32_t)v;
}

// twig of each leaf: first neighbour live at the start of the round
__global__ __launch_bounds__(kT) void k_core_twigs(const uint32_t* __restrict__ leaves,
                                                  int64_t nl,
                                                  const int64_t* __restrict__ row_off,
                                                  const uint32_t* __restrict__ col,
                                                  const uint8_t* __restrict__ alive,
                                                  uint32_t* __restrict__ twig) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nl) return;
  const uint32_t s = leaves[i];
  uint32_t t = kNoParent;
  for (int64_t e = row_off[s]; e < row_off[s + 1]; ++e)
    if (alive[col[e]]) {
      t = col[e];
      break;
    }
  twig[i] = t;
}

__global__ __launch_bounds__(kT) void k_core_by_source(const uint32_t* __restrict__ leaves,
                                                      const uint32_t* __restrict__ twig,
                                                      int64_t nl, uint32_t* __restrict__ deg,
                                                      uint8_t* __restrict__ alive,
                                                      uint32_t* __restrict__ par) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nl || twig[i] == kNoParent) return;
  const uint32_t s = leaves[i];
  deg[s] = 0;
  par[s] = twig[i];
  alive[s] = 0;
}

// A target loses one degree per twig; targets zeroed as leaves this round
// are skipped (a twig can only reach a zero-degree target when both ends
// were leaves, so the decrements never underflow).
__global__ __launch_bounds__(kT) void k_core_by_target(const uint32_t* __restrict__ twig,
                                                      int64_t nl, uint32_t* __restrict__ deg,
                                                      uint8_t* __restrict__ alive) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nl) return;
  const uint32_t t = twig[i];
  if (t == kNoParent || deg[t] == 0) return;
  if (atomicSub(&deg[t], 1u) == 1u) alive[t] = 0;
}

// live out-degree in the core, per vertex
__global__ __launch_bounds__(kT) void k_core_count(const int64_t* __restrict__ row_off,
                                                  const uint32_t* __restrict__ col,
                                                  const uint8_t* __restrict__ alive,
                                                  int64_t n_c, uint64_t* __restrict__ cnt) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n_c) return;
  uint64_t c = 0;
  if (alive[v])
    for (int64_t e = row_off[v]; e < row_off[v + 1]; ++e) c += alive[col[e]];
  cnt[v] = c;
}

// core edge list in original IDs (both directions, already deduplicated)
__global__ __launch_bounds__(kT) void k_core_edges(const int64_t* __restrict__ row_off,
                                                  const uint32_t* __restrict__ col,
                                                  const uint8_t* __restrict__ alive,
                                                  const uint32_t* __restrict__ n2o,
                                                  const uint64_t* __restrict__ off,
                                                  int64_t n_c, uint2* __restrict__ out) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v >= n_c || !alive[v]) return;
  uint64_t p = off[v];
  const uint32_t ov = n2o[v];
  for (int64_t e = row_off[v]; e < row_off[v + 1]; ++e) {
    const uint32_t u = col[e];
    if (alive[u]) out[p++] = make_uint2(ov, n2o[u]);
  }
}

// forest parents by original ID + the has-parent bitmap
__global__ __launch_bounds__(kT) void k_core_forest(const uint32_t* __restrict__ par,
                                                   const uint32_t* __restrict__ n2o,
                                                   int64_t n_c, uint32_t* __restrict__ tp,
                                                   unsigned long long* __restrict__ hp,
                                                   unsigned long long* cnt) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  uint64_t c = 0;
  if (v < n_c && par[v] != kNoParent) {
    const uint32_t o = n2o[v];
    tp[o] = n2o[par[v]];
    atomicOr(&hp[o >> 6], 1ull << (o & 63));
    c = 1;
  }
  c = wave_sum64(c);
  if (lane_id() == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

// reach[o] = o is a core vertex the hd_root BFS visited
__global__ __launch_bounds__(kT) void k_reach_init(const uint32_t* __restrict__ o2n,
                                                  const uint64_t* __restrict__ vis,
                                                  int64_t n, int64_t n_c,
                                                  uint8_t* __restrict__ reach) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (o >= n) return;
  const uint32_t v = o2n[o];
  reach[o] = vis && ((int64_t)v < n_c) && ((vis[v >> 6] >> (v & 63)) & 1ull);
}

// one sweep of "reachable parent -> reachable child"; cnt += changes
__global__ __launch_bounds__(kT) void k_reach_sweep(const uint32_t* __restrict__ tp,
                                                   int64_t n, uint8_t* reach,
                                                   unsigned long long* cnt) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  uint64_t c = 0;
  if (o < n && !reach[o]) {
    const uint32_t p = tp[o];
    if (p != kNoParent && __atomic_load_n(&reach[p], __ATOMIC_RELAXED)) {
      reach[o] = 1;
      c = 1;
    }
  }
  c = wave_sum64(c);
  if (lane_id() == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

__global__ __launch_bounds__(kT) void k_reach_reset(uint32_t* __restrict__ tp,
                                                   const uint8_t* __restrict__ reach,
                                                   int64_t n, unsigned long long* cnt) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  uint64_t c = 0;
  if (o < n && tp[o] != kNoParent && !reach[o]) {
    tp[o] = kNoParent;
    c = 1;
  }
  c = wave_sum64(c);
  if (lane_id() == 0 && c) atomicAdd(cnt, (unsigned long long)c);
}

// largest original ID that is a core vertex or has a parent
__global__ __launch_bounds__(kT) void k_max_used(const uint32_t* __restrict__ o2n,
                                                int64_t n, int64_t n_c,
                                                const uint64_t* __restrict__ hp,
                                                unsigned long long* mx) {
  const int64_t o = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (o >= n) return;
  if ((int64_t)o2n[o] < n_c || ((hp[o >> 6] >> (o & 63)) & 1ull))
    atomicMax(mx, (unsigned long long)o);
}

unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, kT)); }

uint64_t read_u64(const unsigned long long* d, hipStream_t s) {
  unsigned long long h = 0;
  G500_CHECK_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace

void corebfs_prune(g500_graph* g) {
  hipStream_t s = g->stream;
  // 1. full deduplicated symmetric CSR of the input
  build_csr(g, g->edges.get(), g->nedges, g->symmetric_input, -1);
  const int64_t n = g->n, n_c = g->n_c;
  const int log_n = g->log_n;

  // 2. peel leaves
  DevBuf<uint32_t> deg, par, leaves, twig;
  DevBuf<uint8_t> alive;
  DevBuf<unsigned long long> cnt;
  const int64_t nn = std::max<int64_t>(n_c, 1);
  deg.alloc((size_t)nn);
  par.alloc((size_t)nn);
  leaves.alloc((size_t)nn);
  twig.alloc((size_t)nn);
  alive.alloc((size_t)nn);
  cnt.alloc(4);
  if (n_c > 0) {
    hipLaunchKernelGGL(k_core_init, dim3(blocks(n_c)), dim3(kT), 0, s, g->row_off.get(),
                       n_c, deg.get(), alive.get(), par.get());
    G500_CHECK_LAUNCH();
  }
  for (int iter = 0; n_c > 0; ++iter) {
    // Kronecker graphs have a small diameter; the reference gives up past 32
    if (iter > 64) fail(G500_ERR_STATE, "CoreBFS pruning did not converge");
    G500_CHECK_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_core_leaves, dim3(blocks(n_c)), dim3(kT), 0, s, deg.get(), n_c,
                       leaves.get(), cnt.get());
    G500_CHECK_LAUNCH();
    const int64_t nl = (int64_t)read_u64(cnt.get(), s);
    if (nl == 0) break;
    hipLaunchKernelGGL(k_core_twigs, dim3(blocks(nl)), dim3(kT), 0, s, leaves.get(), nl,
                       g->row_off.get(), g->col.get(), alive.get(), twig.get());
    G500_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_core_by_source, dim3(blocks(nl)), dim3(kT), 0, s, leaves.get(),
                       twig.get(), nl, deg.get(), alive.get(), par.get());
    G500_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_core_by_target, dim3(blocks(nl)), dim3(kT), 0, s, twig.get(), nl,
                       deg.get(), alive.get());
    G500_CHECK_LAUNCH();
  }
  deg.release();
  leaves.release();
  twig.release();

  // 3. forest parents by original ID
  g->tree_parent.alloc((size_t)n);
  g->has_parent.alloc((size_t)ceil_div(n, 64));
  G500_CHECK_HIP(hipMemsetAsync(g->tree_parent.get(), 0xFF, n * sizeof(uint32_t), s));
  G500_CHECK_HIP(hipMemsetAsync(g->has_parent.get(), 0, ceil_div(n, 64) * sizeof(uint64_t), s));
  G500_CHECK_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s));
  if (n_c > 0) {
    hipLaunchKernelGGL(k_core_forest, dim3(blocks(n_c)), dim3(kT), 0, s, par.get(),
                       g->new2orig.get(), n_c, g->tree_parent.get(),
                       (unsigned long long*)g->has_parent.get(), cnt.get());
    G500_CHECK_LAUNCH();
  }
  g->tree_vertices = (int64_t)read_u64(cnt.get(), s);
  par.release();

  // 4. core edge list (original IDs) -> core CSR in the same vertex space
  DevBuf<uint64_t> ccnt;
  ccnt.alloc((size_t)nn + 1);
  if (n_c > 0) {
    hipLaunchKernelGGL(k_core_count, dim3(blocks(n_c)), dim3(kT), 0, s, g->row_off.get(),
                       g->col.get(), alive.get(), n_c, ccnt.get());
    G500_CHECK_LAUNCH();
  }
  exclusive_scan<uint64_t>(ccnt.get(), ccnt.get(), n_c, g->scan_ws, s);
  uint64_t m_core = 0;
  G500_CHECK_HIP(hipMemcpyAsync(&m_core, ccnt.get() + n_c, sizeof(uint64_t),
                                hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
  DevBuf<uint2> core_edges;
  core_edges.alloc((size_t)std::max<uint64_t>(m_core, 1));
  if (n_c > 0) {
    hipLaunchKernelGGL(k_core_edges, dim3(blocks(n_c)), dim3(kT), 0, s, g->row_off.get(),
                       g->col.get(), alive.get(), g->new2orig.get(), ccnt.get(), n_c,
                       core_edges.get());
    G500_CHECK_LAUNCH();
  }
  G500_CHECK_HIP(hipStreamSynchronize(s));
  alive.release();
  ccnt.release();
  // drop the full CSR before building the core one
  g->col.release();
  g->first.release();
  g->multi.release();
  g->row_off.release();
  build_csr(g, core_edges.get(), (int64_t)m_core, true, log_n);
  core_edges.release();
  g->core_vertices = g->n_c;

  // find_max_used_vertex over has_edge = core vertex or forest vertex
  G500_CHECK_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_max_used, dim3(blocks(n)), dim3(kT), 0, s, g->orig2new.get(), n,
                     g->n_c, g->has_parent.get(), cnt.get());
  G500_CHECK_LAUNCH();
  g->max_used = (int64_t)read_u64(cnt.get(), s);

  g->h_has_parent.resize((size_t)ceil_div(n, 64));
  G500_CHECK_HIP(hipMemcpy(g->h_has_parent.data(), g->has_parent.get(),
                           g->h_has_parent.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
  g->corebfs = 1;
  g->core_reset_done = false;
  g->hd_root = g->n_c > 0 ? -2 : -1;  // resolved after the new2orig copy below
  if (g->n_c > 0) {
    uint32_t r = 0;
    G500_CHECK_HIP(hipMemcpy(&r, g->new2orig.get(), sizeof(uint32_t), hipMemcpyDeviceToHost));
    g->hd_root = r;  // new ID 0: highest core degree, ties by lowest original ID
  }
  g->constructed = true;
}

void corebfs_reset_unreachable(g500_graph* g, const uint64_t* visited, hipStream_t s) {
  const int64_t n = g->n;
  DevBuf<uint8_t> reach;
  DevBuf<unsigned long long> cnt;
  reach.alloc((size_t)n);
  cnt.alloc(1);
  hipLaunchKernelGGL(k_reach_init, dim3(blocks(n)), dim3(kT), 0, s, g->orig2new.get(),
                     visited, n, g->n_c, reach.get());
  G500_CHECK_LAUNCH();
  for (int iter = 0;; ++iter) {
    if (iter > (1 << 16)) fail(G500_ERR_STATE, "forest reachability did not converge");
    G500_CHECK_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_reach_sweep, dim3(blocks(n)), dim3(kT), 0, s,
                       g->tree_parent.get(), n, reach.get(), cnt.get());
    G500_CHECK_LAUNCH();
    if (read_u64(cnt.get(), s) == 0) break;
  }
  G500_CHECK_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_reach_reset, dim3(blocks(n)), dim3(kT), 0, s, g->tree_parent.get(),
                     reach.get(), n, cnt.get());
  G500_CHECK_LAUNCH();
  g->removed_parents = (int64_t)read_u64(cnt.get(), s);
  g->h_tree_parent.resize((size_t)n);
  G500_CHECK_HIP(hipMemcpy(g->h_tree_parent.data(), g->tree_parent.get(),
                           n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  g->core_reset_done = true;
}

}  // namespace g500
