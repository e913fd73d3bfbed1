// primitives.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Scan: 256-thread blocks own 2048-element tiles; tiles are reduced, the tile
// sums are scanned recursively, then each tile is scanned with its offset.
//
// Radix sort: 8-bit digits, 8192-key tiles (256 threads x 32 keys).  Per pass
//   upsweep   - per-tile digit histogram, counted with wave64 peer masks
//               (8 ballots give every lane the set of lanes sharing its
//               digit, so one lane per digit group adds to LDS),
//   scan      - digit-major [digit][tile] exclusive scan -> global offsets,
//   downsweep - stable in-tile ranks from the same peer masks, keys staged in
//               LDS in digit order, then written out in contiguous runs.
#include "primitives.hpp"

namespace g500 {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t* wsum,
                                           uint64_t* total) {
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
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
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    const uint64_t s = wsum[w];
    off += (w < wave) ? s : 0;
    tot += s;
  }
  __syncthreads();
  if (total) *total = tot;
  return x - v + off;
}

template <typename InT>
__global__ __launch_bounds__(kScanThreads) void k_tile_sums(const InT* in,
                                                           int64_t n,
                                                           uint64_t* sums) {
  __shared__ uint64_t wsum[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t tot;
  (void)block_excl_scan(s, wsum, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Scans each tile; adds tile_off[tile] when given; writes out[n] = total.
template <typename InT>
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(
    const InT* in, uint64_t* out, int64_t n, const uint64_t* tile_off) {
  __shared__ uint64_t tile[kScanTile];
  __shared__ uint64_t wsum[kScanThreads / kWave];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + t;
    tile[k * kScanThreads + t] = (i < n) ? (uint64_t)in[i] : 0;
  }
  __syncthreads();
  uint64_t v[kScanItems];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = tile[t * kScanItems + k];
    s += v[k];
  }
  uint64_t tot;
  uint64_t run = block_excl_scan(s, wsum, &tot);
  const uint64_t off = tile_off ? tile_off[blockIdx.x] : 0;
  run += off;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    tile[t * kScanItems + k] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + t;
    if (i < n) out[i] = tile[k * kScanThreads + t];
  }
  if (t == 0 && base + kScanTile >= n) out[n] = off + tot;
}

int64_t scan_ws_need(int64_t n) {
  int64_t need = 0;
  while (n > kScanTile) {
    const int64_t nt = ceil_div(n, kScanTile);
    need += nt + 1;
    n = nt;
  }
  return need;
}

template <typename InT>
void scan_rec(const InT* in, uint64_t* out, int64_t n, uint64_t* ws,
              hipStream_t s) {
  if (n <= kScanTile) {
    hipLaunchKernelGGL(k_tile_scan<InT>, dim3(1), dim3(kScanThreads), 0, s, in,
                       out, n, (const uint64_t*)nullptr);
    G500_CHECK_LAUNCH();
    return;
  }
  const int64_t nt = ceil_div(n, kScanTile);
  uint64_t* sums = ws;
  hipLaunchKernelGGL(k_tile_sums<InT>, dim3((unsigned)nt), dim3(kScanThreads),
                     0, s, in, n, sums);
  G500_CHECK_LAUNCH();
  scan_rec<uint64_t>(sums, sums, nt, ws + nt + 1, s);
  hipLaunchKernelGGL(k_tile_scan<InT>, dim3((unsigned)nt), dim3(kScanThreads),
                     0, s, in, out, n, (const uint64_t*)sums);
  G500_CHECK_LAUNCH();
}

// ---------------------------------------------------------------- radix sort
// Lanes (among `valid` ones) whose digit equals this lane's digit.
__device__ inline uint64_t peer_mask(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

__global__ __launch_bounds__(kRadixThreads) void k_rs_upsweep(
    const uint64_t* __restrict__ keys, int64_t n, int shift,
    uint32_t* __restrict__ counts, int64_t ntiles) {
  __shared__ uint32_t hist[256];
  const int t = threadIdx.x;
  hist[t] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  for (int r = 0; r < kRadixItems; ++r) {
    const int64_t i = base + r * kRadixThreads + t;
    const bool valid = i < n;
    const uint32_t d = valid ? (uint32_t)(keys[i] >> shift) & 255u : 0u;
    const uint64_t peers = peer_mask(d, valid);
    if (valid && (peers & lanemask_lt()) == 0)
      atomicAdd(&hist[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  counts[(int64_t)t * ntiles + blockIdx.x] = hist[t];
}

__global__ __launch_bounds__(kRadixThreads) void k_rs_downsweep(
    const uint64_t* __restrict__ in, uint64_t* __restrict__ out, int64_t n,
    int shift, const uint32_t* __restrict__ counts,
    const uint64_t* __restrict__ offsets, int64_t ntiles) {
  __shared__ uint64_t stage[kRadixTile];
  __shared__ uint64_t gbase[256];
  __shared__ uint32_t lstart[256];
  __shared__ uint32_t running[256];
  __shared__ uint32_t wcnt[kRadixThreads / kWave][256];
  __shared__ uint64_t wsum[kRadixThreads / kWave];

  const int t = threadIdx.x;
  const int wave = t / kWave;
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kRadixTile;
  const int tile_n = (int)((n - base) < kRadixTile ? (n - base) : kRadixTile);

  const uint32_t c = counts[(int64_t)t * ntiles + tile];
  gbase[t] = offsets[(int64_t)t * ntiles + tile];
  const uint64_t ls = block_excl_scan(c, wsum, nullptr);
  lstart[t] = (uint32_t)ls;
  running[t] = 0;
#pragma unroll
  for (int w = 0; w < kRadixThreads / kWave; ++w) wcnt[w][t] = 0;
  __syncthreads();

  for (int r = 0; r < kRadixItems; ++r) {
    const int li = r * kRadixThreads + t;
    const bool valid = li < tile_n;
    const uint64_t key = valid ? in[base + li] : 0;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const uint64_t peers = peer_mask(d, valid);
    const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = lstart[d] + running[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      stage[pos] = key;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kRadixThreads / kWave; ++w) {
      add += wcnt[w][t];
      wcnt[w][t] = 0;
    }
    running[t] += add;
    __syncthreads();
  }
  for (int j = t; j < tile_n; j += kRadixThreads) {
    const uint64_t key = stage[j];
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    out[gbase[d] + (uint64_t)(j - (int)lstart[d])] = key;
  }
}

}  // namespace

void ScanWorkspace::reserve(int64_t n) {
  const int64_t need = scan_ws_need(n) + 1;
  if (need <= cap) return;
  buf.alloc((size_t)need);
  cap = need;
}

template <typename InT>
void exclusive_scan(const InT* in, uint64_t* out, int64_t n, ScanWorkspace& ws,
                    hipStream_t s) {
  if (n < 0) fail(G500_ERR_ARG, "exclusive_scan: negative length");
  ws.reserve(n);
  scan_rec<InT>(in, out, n, ws.buf.get(), s);
}

template void exclusive_scan<uint32_t>(const uint32_t*, uint64_t*, int64_t,
                                       ScanWorkspace&, hipStream_t);
template void exclusive_scan<uint64_t>(const uint64_t*, uint64_t*, int64_t,
                                       ScanWorkspace&, hipStream_t);

void RadixWorkspace::reserve(int64_t n) {
  if (n <= cap) return;
  const int64_t tiles = ceil_div(n, kRadixTile);
  counts.alloc((size_t)(256 * tiles));
  offsets.alloc((size_t)(256 * tiles + 1));
  scan.reserve(256 * tiles);
  cap = n;
}

uint64_t* radix_sort_u64(uint64_t* keys, uint64_t* alt, int64_t n, int bits,
                         RadixWorkspace& ws, hipStream_t s) {
  if (n <= 1 || bits <= 0) return keys;
  ws.reserve(n);
  const int64_t tiles = ceil_div(n, kRadixTile);
  if (tiles > 0x7FFFFFFF) fail(G500_ERR_ARG, "radix sort: too many keys");
  uint64_t* src = keys;
  uint64_t* dst = alt;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_rs_upsweep, dim3((unsigned)tiles), dim3(kRadixThreads),
                       0, s, src, n, shift, ws.counts.get(), tiles);
    G500_CHECK_LAUNCH();
    exclusive_scan<uint32_t>(ws.counts.get(), ws.offsets.get(), 256 * tiles,
                             ws.scan, s);
    hipLaunchKernelGGL(k_rs_downsweep, dim3((unsigned)tiles),
                       dim3(kRadixThreads), 0, s, src, dst, n, shift,
                       ws.counts.get(), ws.offsets.get(), tiles);
    G500_CHECK_LAUNCH();
    uint64_t* tmp = src;
    src = dst;
    dst = tmp;
  }
  return src;
}

}  // namespace g500
