// gen.hip -- Graph500 Kronecker (R-MAT) edge generation on the GPU, plus the
// host-side MRG table and helpers.
//
// Edge e of the stream starts from the MRG state A^(e * 2^64) * seed
// (generator/graph_generator.hpp:803-806) and draws SCALE quadrants with the
// clip-and-flip rule, then permutes both endpoints with `scramble`
// (graph_generator.hpp:625-637, 835-891).  Each thread produces kEdgesPerThread
// consecutive edges: one table skip for the first, then one multiplication by
// A^(2^64) per following edge.
#include <cstring>
#include <mutex>

#include "internal.hpp"
#include "mrg.hpp"

namespace g500 {

namespace {

void mat5_mul(const Mat5& x, const Mat5& y, Mat5& out) {
  Mat5 r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 5; ++k)
        acc += mrg_reduce((uint64_t)x.a[i * 5 + k] * y.a[k * 5 + j]);
      r.a[i * 5 + j] = mrg_reduce(acc);
    }
  out = r;
}

Mat5 mat5_identity() {
  Mat5 m;
  std::memset(&m, 0, sizeof(m));
  for (int i = 0; i < 5; ++i) m.a[i * 5 + i] = 1;
  return m;
}

std::once_flag g_table_once;
MrgTable* g_table = nullptr;

void build_table() {
  auto* t = new MrgTable;
  Mat5 a;
  std::memset(&a, 0, sizeof(a));
  a.a[0] = kMrgX;
  a.a[4] = kMrgY;
  for (int i = 1; i < 5; ++i) a.a[i * 5 + (i - 1)] = 1;
  Mat5 base = a;  // A^(256^k)
  for (int k = 0; k < 24; ++k) {
    t->t[k][0] = mat5_identity();
    t->t[k][1] = base;
    for (int b = 2; b < 256; ++b) mat5_mul(t->t[k][b - 1], base, t->t[k][b]);
    mat5_mul(t->t[k][255], base, base);  // A^(256^(k+1))
  }
  g_table = t;
}

constexpr int kGenThreads = 256;
constexpr int kEdgesPerThread = 8;

__device__ inline uint64_t scramble_dev(uint64_t v, int scale, uint64_t val0,
                                        uint64_t val1) {
  v += val0 + val1;
  v *= (val0 | 0x4519840211493211ull);
  v = __builtin_bitreverse64(v) >> (64 - scale);
  v *= (val1 | 0x3050852102C843A5ull);
  v = __builtin_bitreverse64(v) >> (64 - scale);
  return v;
}

// Quadrant in [0,4) drawn as 1 (p .19), 2 (p .19), 0 (p .57), 3 (p .05) from
// a rejection-sampled value in [0, 10000).
__device__ inline int quadrant_dev(MrgState& st) {
  constexpr uint32_t kLimit = 0xFFFFFFFFu % 10000u;
  uint32_t val = mrg_next(st);
  while (val < kLimit) val = mrg_next(st);
  val %= 10000u;
  if (val < 1900u) return 1;
  val -= 1900u;
  if (val < 1900u) return 2;
  val -= 1900u;
  return val < 5700u ? 0 : 3;
}

struct GenArgs {
  MrgState seed;
  uint64_t val0, val1;
  int64_t nedges;
  int scale;
};

__global__ __launch_bounds__(kGenThreads) void k_generate(
    const Mat5* __restrict__ mid_table,  // byte positions 8..15, [8][256]
    GenArgs a, uint2* __restrict__ out) {
  const int64_t e0 =
      ((int64_t)blockIdx.x * kGenThreads + threadIdx.x) * kEdgesPerThread;
  if (e0 >= a.nedges) return;
  MrgState cur = a.seed;
#pragma unroll 1
  for (int k = 0; k < 8; ++k) {
    const uint32_t b = (uint32_t)(((uint64_t)e0 >> (8 * k)) & 0xFF);
    if (b) mat5_apply(mid_table[k * 256 + b], cur);
  }
  const Mat5 step = mid_table[1];  // A^(2^64)
  const int64_t n_top = (int64_t)1 << a.scale;
#pragma unroll 1
  for (int j = 0; j < kEdgesPerThread; ++j) {
    const int64_t e = e0 + j;
    if (e >= a.nedges) break;
    MrgState st = cur;
    int64_t n = n_top, src = 0, tgt = 0;
    while (n > 1) {
      const int q = quadrant_dev(st);
      int so = q >> 1, to = q & 1;
      if (src == tgt && so > to) {
        so = 0;
        to = 1;
      }
      n >>= 1;
      src += n * so;
      tgt += n * to;
    }
    out[e] = make_uint2((uint32_t)scramble_dev((uint64_t)src, a.scale, a.val0, a.val1),
                        (uint32_t)scramble_dev((uint64_t)tgt, a.scale, a.val0, a.val1));
    if (j + 1 < kEdgesPerThread) mat5_apply(step, cur);
  }
}

}  // namespace

const MrgTable& mrg_table() {
  std::call_once(g_table_once, build_table);
  return *g_table;
}

void mrg_skip_host(MrgState& s, uint64_t hi, uint64_t mid, uint64_t lo) {
  const MrgTable& t = mrg_table();
  const uint64_t w[3] = {lo, mid, hi};
  for (int word = 0; word < 3; ++word)
    for (int byte = 0; byte < 8; ++byte) {
      const uint32_t b = (uint32_t)((w[word] >> (8 * byte)) & 0xFF);
      if (b) mat5_apply(t.t[word * 8 + byte][b], s);
    }
}

double mrg_get_double_host(MrgState& s) {
  const double hi = (double)mrg_next(s);
  const double lo = (double)mrg_next(s);
  return hi * .000000000465661287524579692 +
         lo * .0000000000000000002168404346990492787;
}

void make_random_numbers_host(int64_t n, uint64_t s1, uint64_t s2,
                              int64_t position, double* out) {
  MrgState st = mrg_seed_from_user(s1, s2);
  mrg_skip_host(st, 2, 0, 2 * (uint64_t)position);
  for (int64_t i = 0; i < n; ++i) out[i] = mrg_get_double_host(st);
}

void scramble_values_host(uint64_t s1, uint64_t s2, uint64_t out[2]) {
  MrgState st = mrg_seed_from_user(s1, s2);
  mrg_skip_host(st, 50, 7, 0);
  uint64_t a = mrg_next(st);
  a = a * 0xFFFFFFFFull + mrg_next(st);
  uint64_t b = mrg_next(st);
  b = b * 0xFFFFFFFFull + mrg_next(st);
  out[0] = a;
  out[1] = b;
}

void generate_edges(g500_graph* g, int scale, int edgefactor, uint64_t seed1,
                    uint64_t seed2) {
  if (scale < 1 || scale > 32)
    fail(G500_ERR_ARG, "scale %d outside [1, 32]", scale);
  if (edgefactor < 1) fail(G500_ERR_ARG, "edgefactor %d < 1", edgefactor);
  const int64_t ne = ((int64_t)1 << scale) * edgefactor;
  const MrgTable& t = mrg_table();
  DevBuf<Mat5> tab;
  tab.alloc(8 * 256);
  G500_CHECK_HIP(hipMemcpyAsync(tab.get(), &t.t[8][0], sizeof(Mat5) * 8 * 256,
                                hipMemcpyHostToDevice, g->stream));
  GenArgs a;
  a.seed = mrg_seed_from_user(seed1, seed2);
  uint64_t sv[2];
  scramble_values_host(seed1, seed2, sv);
  a.val0 = sv[0];
  a.val1 = sv[1];
  a.nedges = ne;
  a.scale = scale;
  g->edges.alloc((size_t)ne);
  g->nedges = ne;
  g->id_bits = scale;
  const int64_t threads = ceil_div(ne, kEdgesPerThread);
  const int64_t blocks = ceil_div(threads, kGenThreads);
  hipLaunchKernelGGL(k_generate, dim3((unsigned)blocks), dim3(kGenThreads), 0,
                     g->stream, tab.get(), a, g->edges.get());
  G500_CHECK_LAUNCH();
  G500_CHECK_HIP(hipStreamSynchronize(g->stream));
}

}  // namespace g500
