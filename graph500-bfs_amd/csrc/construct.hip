// construct.hip -- symmetrize, degree-descending relabel and CSR build on GPU.
//
// Semantics follow the reference's single-rank construction:
//  - self loops dropped, both directions kept, vertex space = 2^(msb(OR of
//    endpoint IDs)+1)          (mpi/benchmark_helper.hpp:389-497)
//  - degree = symmetric edge count including duplicates, vertices renumbered
//    by descending degree      (mpi/graph_constructor.hpp:572-761)
//  - each row sorted by renumbered target, duplicates merged
//    (mpi/graph_constructor.hpp:1968-2101)
// Ties in degree are broken by ascending original ID (the reference's
// std::sort leaves them unspecified), which makes the whole CSR a pure
// function of the edge list.
//
// Pipeline: atomics for degrees -> radix sort of (inverted degree, ID) keys ->
// relabel maps -> (new_src, new_dst) keys, loops as a max-value sentinel ->
// radix sort -> flag/scan/scatter dedupe into row offsets + uint32 columns.
#include <algorithm>

#include "internal.hpp"

namespace g500 {

namespace {

constexpr int kT = 256;

__device__ inline uint64_t wave_or(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ inline uint64_t wave_max(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = __shfl_xor(v, o, kWave);
    v = w > v ? w : v;
  }
  return v;
}

// stats[0] |= endpoint IDs of non-loop edges, stats[1] += non-loop edges
// (symmetric input: each directed edge counts once, toward its source)
__global__ __launch_bounds__(kT) void k_degree(const uint2* __restrict__ e,
                                              int64_t ne, int symmetric,
                                              uint32_t* __restrict__ deg,
                                              unsigned long long* stats) {
  uint64_t orv = 0, cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < ne;
       i += (int64_t)gridDim.x * kT) {
    const uint2 x = e[i];
    if (x.x == x.y) continue;
    atomicAdd(&deg[x.x], 1u);
    if (!symmetric) atomicAdd(&deg[x.y], 1u);
    orv |= (uint64_t)(x.x | x.y);
    ++cnt;
  }
  orv = wave_or(orv);
  cnt = wave_sum(cnt);
  if (lane_id() == 0) {
    if (orv) atomicOr(&stats[0], (unsigned long long)orv);
    if (cnt) atomicAdd(&stats[1], (unsigned long long)cnt);
  }
}

// stats[2] = max degree, stats[3] = #vertices with degree > 0,
// stats[4] = max vertex with degree > 0
__global__ __launch_bounds__(kT) void k_deg_stats(const uint32_t* __restrict__ deg,
                                                 int64_t n,
                                                 unsigned long long* stats) {
  uint64_t mx = 0, nz = 0, mu = 0;
  for (int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * kT) {
    const uint32_t d = deg[v];
    mx = d > mx ? d : mx;
    if (d) {
      ++nz;
      mu = (uint64_t)v;
    }
  }
  mx = wave_max(mx);
  nz = wave_sum(nz);
  mu = wave_max(mu);
  if (lane_id() == 0) {
    if (mx) atomicMax(&stats[2], (unsigned long long)mx);
    if (nz) atomicAdd(&stats[3], (unsigned long long)nz);
    if (mu) atomicMax(&stats[4], (unsigned long long)mu);
  }
}

__global__ __launch_bounds__(kT) void k_vertex_keys(const uint32_t* __restrict__ deg,
                                                   int64_t n, int lg,
                                                   uint64_t maxdeg,
                                                   uint64_t* __restrict__ keys) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n) keys[v] = ((maxdeg - deg[v]) << lg) | (uint64_t)v;
}

__global__ __launch_bounds__(kT) void k_relabel(const uint64_t* __restrict__ keys,
                                               int64_t n, uint64_t mask,
                                               const uint32_t* __restrict__ deg,
                                               uint32_t* __restrict__ n2o,
                                               uint32_t* __restrict__ o2n,
                                               uint32_t* __restrict__ deg_new) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = (uint32_t)(keys[i] & mask);
  n2o[i] = o;
  o2n[o] = (uint32_t)i;
  deg_new[i] = deg[o];
}

// Two keys per input edge ((a,b) and (b,a)), or one when the input is
// already symmetric; self loops become the sentinel (sorts last).
__global__ __launch_bounds__(kT) void k_sym_keys(const uint2* __restrict__ e,
                                                int64_t ne, int symmetric,
                                                const uint32_t* __restrict__ o2n,
                                                int lg, uint64_t sentinel,
                                                uint64_t* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= ne) return;
  const uint2 x = e[i];
  uint64_t k0 = sentinel, k1 = sentinel;
  if (x.x != x.y) {
    const uint64_t a = o2n[x.x], b = o2n[x.y];
    k0 = (a << lg) | b;
    k1 = (b << lg) | a;
  }
  if (symmetric) {
    keys[i] = k0;
  } else {
    keys[2 * i] = k0;
    keys[2 * i + 1] = k1;
  }
}

constexpr int kDedupItems = 16;
constexpr int kDedupTile = kT * kDedupItems;

__device__ inline uint64_t block_excl_scan_u64(uint64_t v, uint64_t* wsum,
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
  for (int w = 0; w < kT / kWave; ++w) {
    off += (w < wave) ? wsum[w] : 0;
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return x - v + off;
}

// Per tile: number of keys that differ from their predecessor.
__global__ __launch_bounds__(kT) void k_dedup_count(const uint64_t* __restrict__ k,
                                                   int64_t n,
                                                   uint32_t* __restrict__ cnt) {
  __shared__ uint64_t wsum[kT / kWave];
  const int64_t base = (int64_t)blockIdx.x * kDedupTile + threadIdx.x * kDedupItems;
  uint32_t c = 0;
  for (int j = 0; j < kDedupItems; ++j) {
    const int64_t i = base + j;
    if (i < n && (i == 0 || k[i] != k[i - 1])) ++c;
  }
  uint64_t tot;
  (void)block_excl_scan_u64(c, wsum, &tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = (uint32_t)tot;
}

__global__ __launch_bounds__(kT) void k_dedup_write(
    const uint64_t* __restrict__ k, int64_t n, int lg,
    const uint64_t* __restrict__ tile_off, uint32_t* __restrict__ col,
    int64_t* __restrict__ row_off) {
  __shared__ uint64_t wsum[kT / kWave];
  const int64_t base = (int64_t)blockIdx.x * kDedupTile + threadIdx.x * kDedupItems;
  const uint64_t mask = (1ull << lg) - 1;
  uint32_t c = 0;
  for (int j = 0; j < kDedupItems; ++j) {
    const int64_t i = base + j;
    if (i < n && (i == 0 || k[i] != k[i - 1])) ++c;
  }
  uint64_t tot;
  uint64_t pos = tile_off[blockIdx.x] + block_excl_scan_u64(c, wsum, &tot);
  for (int j = 0; j < kDedupItems; ++j) {
    const int64_t i = base + j;
    if (i >= n) break;
    const uint64_t cur = k[i];
    const uint64_t prev = i ? k[i - 1] : ~cur;
    if (cur == prev) continue;
    col[pos] = (uint32_t)(cur & mask);
    if (i == 0 || (cur >> lg) != (prev >> lg)) row_off[cur >> lg] = (int64_t)pos;
    ++pos;
  }
}

__global__ void k_set_i64(int64_t* p, int64_t v) { *p = v; }

// Copy of each row's first entry with its original ID, stored densely so that
// the bottom-up sweep reads both coalesced (the reference keeps the same
// "isolated first edge", mpi/graph_constructor.hpp:1387-1557).
__global__ __launch_bounds__(kT) void k_first_nbr(const int64_t* __restrict__ row_off,
                                                 const uint32_t* __restrict__ col,
                                                 const uint32_t* __restrict__ n2o,
                                                 int64_t n_c, uint2* __restrict__ first) {
  const int64_t v = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (v < n_c) {
    const uint32_t u = col[row_off[v]];
    first[v] = make_uint2(u, n2o[u]);
  }
}

// Bit v: row v has more than one entry (a bottom-up vertex whose first
// neighbour is unvisited has nothing else to scan otherwise).
__global__ __launch_bounds__(kT) void k_multi_bits(const int64_t* __restrict__ row_off,
                                                  int64_t n_c, uint64_t* __restrict__ multi) {
  const int64_t w = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (w * 64 >= n_c) return;
  uint64_t bits = 0;
  const int64_t v0 = w * 64;
  const int cnt = (int)(n_c - v0 < 64 ? n_c - v0 : 64);
  for (int b = 0; b < cnt; ++b)
    if (row_off[v0 + b + 1] - row_off[v0 + b] > 1) bits |= 1ull << b;
  multi[w] = bits;
}

int bitlen(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

}  // namespace

void build_csr(g500_graph* g, const uint2* edges, int64_t ne, bool symmetric,
               int force_log_n) {
  hipStream_t s = g->stream;
  const int64_t nv = (int64_t)1 << g->id_bits;  // covers every input ID

  DevBuf<unsigned long long> stats;
  stats.alloc(8);
  G500_CHECK_HIP(hipMemsetAsync(stats.get(), 0, 8 * sizeof(unsigned long long), s));
  DevBuf<uint32_t> deg;
  deg.alloc((size_t)nv);
  G500_CHECK_HIP(hipMemsetAsync(deg.get(), 0, nv * sizeof(uint32_t), s));
  if (ne > 0) {
    const int64_t blocks = std::min<int64_t>(ceil_div(ne, kT), 8192);
    hipLaunchKernelGGL(k_degree, dim3((unsigned)blocks), dim3(kT), 0, s,
                       edges, ne, (int)symmetric, deg.get(), stats.get());
    G500_CHECK_LAUNCH();
  }
  {
    const int64_t blocks = std::min<int64_t>(ceil_div(nv, kT), 8192);
    hipLaunchKernelGGL(k_deg_stats, dim3((unsigned)blocks), dim3(kT), 0, s,
                       deg.get(), nv, stats.get());
    G500_CHECK_LAUNCH();
  }
  unsigned long long h[8];
  G500_CHECK_HIP(hipMemcpyAsync(h, stats.get(), sizeof(h), hipMemcpyDeviceToHost, s));
  G500_CHECK_HIP(hipStreamSynchronize(s));
  const uint64_t ormask = h[0];
  const int64_t nonloop = (int64_t)h[1];
  const uint64_t maxdeg = h[2];
  const int lg = force_log_n >= 0 ? force_log_n : bitlen(ormask);  // msb + 1
  const int64_t n = (int64_t)1 << lg;
  g->log_n = lg;
  g->n = n;
  g->n_c = (int64_t)h[3];
  g->max_used = (int64_t)h[4];
  g->max_deg = (int64_t)maxdeg;
  g->m_sym = symmetric ? nonloop : 2 * nonloop;

  // ---- relabel by (degree desc, id asc)
  g->new2orig.alloc((size_t)n);
  g->orig2new.alloc((size_t)n);
  g->deg_new.alloc((size_t)n);
  {
    DevBuf<uint64_t> vk, vk_alt;
    vk.alloc((size_t)n);
    vk_alt.alloc((size_t)n);
    const int64_t blocks = ceil_div(n, kT);
    hipLaunchKernelGGL(k_vertex_keys, dim3((unsigned)blocks), dim3(kT), 0, s,
                       deg.get(), n, lg, (uint64_t)maxdeg, vk.get());
    G500_CHECK_LAUNCH();
    uint64_t* sorted = radix_sort_u64(vk.get(), vk_alt.get(), n,
                                      lg + std::max(1, bitlen(maxdeg)),
                                      g->radix_ws, s);
    hipLaunchKernelGGL(k_relabel, dim3((unsigned)blocks), dim3(kT), 0, s, sorted,
                       n, (uint64_t)(n - 1), deg.get(), g->new2orig.get(),
                       g->orig2new.get(), g->deg_new.get());
    G500_CHECK_LAUNCH();
    G500_CHECK_HIP(hipStreamSynchronize(s));
  }
  deg.release();

  // ---- symmetric (new_src, new_dst) keys, sorted, then deduplicated
  {
    const int64_t nk = symmetric ? ne : 2 * ne;
    DevBuf<uint64_t> ek, ek_alt;
    ek.alloc((size_t)nk);
    ek_alt.alloc((size_t)nk);
    const uint64_t sentinel = ((uint64_t)(n - 1) << lg) | (uint64_t)(n - 1);
    if (ne > 0) {
      hipLaunchKernelGGL(k_sym_keys, dim3((unsigned)ceil_div(ne, kT)), dim3(kT), 0,
                         s, edges, ne, (int)symmetric, g->orig2new.get(), lg, sentinel,
                         ek.get());
      G500_CHECK_LAUNCH();
    }
    uint64_t* sorted = radix_sort_u64(ek.get(), ek_alt.get(), nk, 2 * lg,
                                      g->radix_ws, s);
    const int64_t ms = g->m_sym;  // sentinels sort to the end
    const int64_t tiles = std::max<int64_t>(1, ceil_div(ms, kDedupTile));
    DevBuf<uint32_t> tcnt;
    DevBuf<uint64_t> toff;
    tcnt.alloc((size_t)tiles);
    toff.alloc((size_t)tiles + 1);
    hipLaunchKernelGGL(k_dedup_count, dim3((unsigned)tiles), dim3(kT), 0, s,
                       sorted, ms, tcnt.get());
    G500_CHECK_LAUNCH();
    exclusive_scan<uint32_t>(tcnt.get(), toff.get(), tiles, g->scan_ws, s);
    uint64_t m_total = 0;
    G500_CHECK_HIP(hipMemcpyAsync(&m_total, toff.get() + tiles, sizeof(uint64_t),
                                  hipMemcpyDeviceToHost, s));
    G500_CHECK_HIP(hipStreamSynchronize(s));
    g->m = (int64_t)m_total;
    // padding: bottom-up scans read aligned 16-byte groups, which may run a
    // few entries past the last row
    g->col.alloc((size_t)g->m + kColPad);
    g->row_off.alloc((size_t)g->n_c + 1);
    hipLaunchKernelGGL(k_dedup_write, dim3((unsigned)tiles), dim3(kT), 0, s,
                       sorted, ms, lg, toff.get(), g->col.get(), g->row_off.get());
    G500_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(1), 0, s, g->row_off.get() + g->n_c,
                       g->m);
    G500_CHECK_LAUNCH();
    g->first.alloc((size_t)std::max<int64_t>(g->n_c, 1));
    g->multi.alloc((size_t)ceil_div(std::max<int64_t>(g->n_c, 1), 64));
    if (g->n_c > 0) {
      hipLaunchKernelGGL(k_first_nbr, dim3((unsigned)ceil_div(g->n_c, kT)), dim3(kT), 0, s,
                         g->row_off.get(), g->col.get(), g->new2orig.get(), g->n_c,
                         g->first.get());
      G500_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_multi_bits, dim3((unsigned)ceil_div(ceil_div(g->n_c, 64), kT)),
                         dim3(kT), 0, s, g->row_off.get(), g->n_c, g->multi.get());
      G500_CHECK_LAUNCH();
    }
    G500_CHECK_HIP(hipStreamSynchronize(s));
  }
  g->h_orig2new.clear();
}

void construct_graph(g500_graph* g, int corebfs) {
  build_csr(g, g->edges.get(), g->nedges, g->symmetric_input, -1);
  g->corebfs = corebfs;
  g->constructed = true;
}

void ensure_host_orig2new(g500_graph* g) {
  if (!g->h_orig2new.empty() || g->n == 0) return;
  g->h_orig2new.resize((size_t)g->n);
  G500_CHECK_HIP(hipMemcpy(g->h_orig2new.data(), g->orig2new.get(),
                           g->n * sizeof(uint32_t), hipMemcpyDeviceToHost));
}

}  // namespace g500

int64_t g500_graph::device_bytes() const {
  return (int64_t)(edges.n * sizeof(uint2) + row_off.n * sizeof(int64_t) +
                   col.n * sizeof(uint32_t) + first.n * sizeof(uint2) +
                   multi.n * sizeof(uint64_t) +
                   (deg_new.n + new2orig.n + orig2new.n) * sizeof(uint32_t) +
                   tree_parent.n * sizeof(uint32_t) + has_parent.n * sizeof(uint64_t));
}
