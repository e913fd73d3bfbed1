/*
 * g500_oracle.c -- CPU restatement of the Graph500-BFS single-rank path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never links this file:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * the library built from it.
 *
 * Written from the reference's behaviour, not its code: the MRG skip uses
 * plain 5x5 matrix powers (A^(2^k) by squaring) instead of the reference's
 * precomputed byte table; the CSR is built with an LSD radix sort over
 * (new_src, new_dst) keys; BFS levels come from a textbook queue BFS.
 *
 * What pins each part (see tests/):
 *   MRG            -> the reference's own splittable_mrg.c compiled into
 *                     oracle/_ref (bit-exact states and doubles)
 *   generator/CSR  -> pf_nedge golden table (mpi/utils.hpp:3688-3733) and the
 *                     edge counts the reference logs (SCALE 16/20/24)
 *   BFS levels     -> scipy.sparse.csgraph.breadth_first_order distances
 *   validator      -> planted mistakes (each check must fire) + valid trees
 *   roots          -> _ref doubles + brute-force restatement in the tests
 *   DO schedule    -> hand-traced level sequences in the tests
 */
#include "g500_oracle.h"

#include <stdlib.h>
#include <string.h>

#define MRG_P 0x7FFFFFFFu
#define MRG_X 107374182u
#define MRG_Y 104480u

/* ------------------------------------------------------------------------ */
/* MRG: state s = (z1..z5); one step s' = A s with                          */
/*   z1' = x*z1 + y*z5, z2' = z1, z3' = z2, z4' = z3, z5' = z4  (mod 2^31-1)  */
/* (generator/splittable_mrg.c:17-31, 210-216).                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint32_t a[5][5];
} mat5;

static uint32_t mulmod(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * b) % MRG_P);
}

static void mat_mul(const mat5* x, const mat5* y, mat5* out) {
  mat5 r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 5; ++k) acc += mulmod(x->a[i][k], y->a[k][j]);
      r.a[i][j] = (uint32_t)(acc % MRG_P);
    }
  *out = r;
}

static void mat_apply(const mat5* m, orc_mrg_state* st) {
  const uint32_t z[5] = {st->z1, st->z2, st->z3, st->z4, st->z5};
  uint32_t o[5];
  for (int i = 0; i < 5; ++i) {
    uint64_t acc = 0;
    for (int k = 0; k < 5; ++k) acc += mulmod(m->a[i][k], z[k]);
    o[i] = (uint32_t)(acc % MRG_P);
  }
  st->z1 = o[0];
  st->z2 = o[1];
  st->z3 = o[2];
  st->z4 = o[3];
  st->z5 = o[4];
}

/* pow2[k] = A^(2^k), k = 0..191 */
static mat5 g_pow2[192];
static int g_pow2_ready = 0;

static void init_pow2(void) {
  if (g_pow2_ready) return;
  mat5 a;
  memset(&a, 0, sizeof(a));
  a.a[0][0] = MRG_X;
  a.a[0][4] = MRG_Y;
  for (int i = 1; i < 5; ++i) a.a[i][i - 1] = 1;
  g_pow2[0] = a;
  for (int k = 1; k < 192; ++k) mat_mul(&g_pow2[k - 1], &g_pow2[k - 1], &g_pow2[k]);
  g_pow2_ready = 1;
}

void orc_make_mrg_seed(uint64_t s1, uint64_t s2, orc_mrg_state* st) {
  st->z1 = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
  st->z2 = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
  st->z3 = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
  st->z4 = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
  st->z5 = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
}

void orc_mrg_skip(orc_mrg_state* st, uint64_t hi, uint64_t mid, uint64_t lo) {
  init_pow2();
  const uint64_t w[3] = {lo, mid, hi};
  for (int word = 0; word < 3; ++word)
    for (int b = 0; b < 64; ++b)
      if ((w[word] >> b) & 1) mat_apply(&g_pow2[word * 64 + b], st);
}

uint32_t orc_mrg_get_uint(orc_mrg_state* st) {
  const uint32_t nz =
      (uint32_t)(((uint64_t)MRG_X * st->z1 + (uint64_t)MRG_Y * st->z5) % MRG_P);
  st->z5 = st->z4;
  st->z4 = st->z3;
  st->z3 = st->z2;
  st->z2 = st->z1;
  st->z1 = nz;
  return nz;
}

double orc_mrg_get_double(orc_mrg_state* st) {
  /* two draws, first one scaled by (2^31-1)^-1, second by (2^31-1)^-2 */
  const double hi = (double)orc_mrg_get_uint(st);
  const double lo = (double)orc_mrg_get_uint(st);
  return hi * .000000000465661287524579692 +
         lo * .0000000000000000002168404346990492787;
}

void orc_make_random_numbers(int64_t n, uint64_t s1, uint64_t s2, int64_t pos,
                             double* out) {
  orc_mrg_state st;
  orc_make_mrg_seed(s1, s2, &st);
  orc_mrg_skip(&st, 2, 0, 2 * (uint64_t)pos); /* two draws per double */
  for (int64_t i = 0; i < n; ++i) out[i] = orc_mrg_get_double(&st);
}

/* ------------------------------------------------------------------------ */
/* Kronecker / R-MAT generator                                               */
/* ------------------------------------------------------------------------ */
static uint64_t bitrev64(uint64_t x) {
  uint64_t r = 0;
  for (int i = 0; i < 64; ++i) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

void orc_scramble_values(uint64_t s1, uint64_t s2, uint64_t out[2]) {
  orc_mrg_state st;
  orc_make_mrg_seed(s1, s2, &st);
  orc_mrg_skip(&st, 50, 7, 0);
  uint64_t a = orc_mrg_get_uint(&st);
  a = a * UINT64_C(0xFFFFFFFF) + orc_mrg_get_uint(&st);
  uint64_t b = orc_mrg_get_uint(&st);
  b = b * UINT64_C(0xFFFFFFFF) + orc_mrg_get_uint(&st);
  out[0] = a;
  out[1] = b;
}

int64_t orc_scramble(int64_t v0, int scale, uint64_t val0, uint64_t val1) {
  uint64_t v = (uint64_t)v0;
  v += val0 + val1;
  v *= (val0 | UINT64_C(0x4519840211493211));
  v = bitrev64(v) >> (64 - scale);
  v *= (val1 | UINT64_C(0x3050852102C843A5));
  v = bitrev64(v) >> (64 - scale);
  return (int64_t)v;
}

/* One quadrant draw in [0,4): rejection-sample a value in [0, 10000) from the
 * MRG output (values below 2^32-1 mod 10000 = 7295 are redrawn), then
 * B = 1 (p .19), C = 2 (p .19), A = 0 (p .57), D = 3 (p .05). */
static int quadrant(orc_mrg_state* st) {
  const uint32_t limit = 0xFFFFFFFFu % 10000u;
  uint32_t val = orc_mrg_get_uint(st);
  while (val < limit) val = orc_mrg_get_uint(st);
  val %= 10000u;
  if (val < 1900u) return 1;
  val -= 1900u;
  if (val < 1900u) return 2;
  val -= 1900u;
  if (val < 5700u) return 0;
  return 3;
}

static void one_edge(int scale, orc_mrg_state* st, uint64_t sv0, uint64_t sv1,
                     int64_t* out0, int64_t* out1) {
  int64_t n = (int64_t)1 << scale;
  int64_t src = 0, tgt = 0;
  while (n > 1) {
    const int q = quadrant(st);
    int so = q / 2, to = q % 2;
    if (src == tgt && so > to) { /* clip-and-flip: keep src <= tgt */
      const int t = so;
      so = to;
      to = t;
    }
    n /= 2;
    src += n * so;
    tgt += n * to;
  }
  *out0 = orc_scramble(src, scale, sv0, sv1);
  *out1 = orc_scramble(tgt, scale, sv0, sv1);
}

void orc_generate_edges(int scale, uint64_t s1, uint64_t s2, int64_t start,
                        int64_t end, int64_t* v0, int64_t* v1) {
  init_pow2();
  uint64_t sv[2];
  orc_scramble_values(s1, s2, sv);
  orc_mrg_state seed;
  orc_make_mrg_seed(s1, s2, &seed);
  /* Edge e starts from A^(e * 2^64) * seed.  Chunks start with a bitwise skip
   * and then advance by one multiplication with A^(2^64) per edge. */
  const int64_t CH = 4096;
  const int64_t nchunks = (end - start + CH - 1) / CH;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t c = 0; c < nchunks; ++c) {
    const int64_t e0 = start + c * CH;
    const int64_t e1 = (e0 + CH < end) ? e0 + CH : end;
    orc_mrg_state cur = seed;
    orc_mrg_skip(&cur, 0, (uint64_t)e0, 0);
    for (int64_t e = e0; e < e1; ++e) {
      orc_mrg_state st = cur;
      one_edge(scale, &st, sv[0], sv[1], &v0[e - start], &v1[e - start]);
      mat_apply(&g_pow2[64], &cur);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Graph construction                                                        */
/* ------------------------------------------------------------------------ */
struct orc_graph {
  int log_n;
  int64_t n, n_c, m, m_sym, max_used;
  int64_t* row_off;   /* n+1 */
  uint32_t* col;      /* m */
  int64_t* deg_new;   /* n */
  uint32_t* new2orig; /* n */
  uint32_t* orig2new; /* n */
  int owns;
};

static int msb64(uint64_t x) { return 63 - __builtin_clzll(x); }

/* LSD radix sort of 64-bit keys over the low `bits` bits, 11-bit digits. */
static int radix_sort_u64(uint64_t* keys, int64_t n, int bits) {
  if (n <= 1 || bits <= 0) return 0;
  uint64_t* tmp = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
  if (!tmp) return -1;
  const int D = 11, R = 1 << D;
  int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * R);
  uint64_t *src = keys, *dst = tmp;
  for (int shift = 0; shift < bits; shift += D) {
    memset(cnt, 0, sizeof(int64_t) * R);
    for (int64_t i = 0; i < n; ++i) cnt[(src[i] >> shift) & (R - 1)]++;
    int64_t s = 0;
    for (int d = 0; d < R; ++d) {
      const int64_t c = cnt[d];
      cnt[d] = s;
      s += c;
    }
    for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & (R - 1)]++] = src[i];
    uint64_t* t = src;
    src = dst;
    dst = t;
  }
  if (src != keys) memcpy(keys, src, (size_t)n * sizeof(uint64_t));
  free(cnt);
  free(tmp);
  return 0;
}

orc_graph* orc_graph_build(const int64_t* v0, const int64_t* v1, int64_t ne) {
  orc_graph* g = (orc_graph*)calloc(1, sizeof(orc_graph));
  if (!g) return NULL;
  g->owns = 1;
  uint64_t ormask = 0;
  int64_t loops = 0;
  for (int64_t i = 0; i < ne; ++i) {
    if (v0[i] == v1[i]) {
      ++loops;
      continue;
    }
    ormask |= (uint64_t)(v0[i] | v1[i]);
  }
  g->log_n = ormask ? msb64(ormask) + 1 : 0;
  const int lg = g->log_n;
  const int64_t n = (int64_t)1 << lg;
  g->n = n;
  int64_t* deg = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  g->deg_new = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  g->new2orig = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  g->orig2new = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  g->row_off = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  if (!deg || !g->deg_new || !g->new2orig || !g->orig2new || !g->row_off) {
    free(deg);
    orc_graph_free(g);
    return NULL;
  }
  for (int64_t i = 0; i < ne; ++i) {
    if (v0[i] == v1[i]) continue;
    deg[v0[i]]++;
    deg[v1[i]]++;
  }
  int64_t maxdeg = 0;
  for (int64_t v = 0; v < n; ++v)
    if (deg[v] > maxdeg) maxdeg = deg[v];
  /* order by (degree desc, id asc): key = ((maxdeg - deg) << lg) | id */
  const int degbits = maxdeg ? msb64((uint64_t)maxdeg) + 1 : 1;
  uint64_t* vk = (uint64_t*)malloc((size_t)n * sizeof(uint64_t));
  if (!vk) {
    free(deg);
    orc_graph_free(g);
    return NULL;
  }
  for (int64_t v = 0; v < n; ++v)
    vk[v] = ((uint64_t)(maxdeg - deg[v]) << lg) | (uint64_t)v;
  radix_sort_u64(vk, n, lg + degbits);
  g->n_c = 0;
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t o = (uint32_t)(vk[i] & (((uint64_t)1 << lg) - 1));
    g->new2orig[i] = o;
    g->orig2new[o] = (uint32_t)i;
    g->deg_new[i] = deg[o];
    if (deg[o] > 0) g->n_c = i + 1;
  }
  free(vk);
  g->max_used = 0;
  for (int64_t v = n - 1; v >= 0; --v)
    if (deg[v] > 0) {
      g->max_used = v;
      break;
    }
  free(deg);
  /* symmetric keys (new_src << lg) | new_dst, then sort + merge duplicates */
  const int64_t ms = 2 * (ne - loops);
  g->m_sym = ms;
  uint64_t* ek = (uint64_t*)malloc((size_t)(ms ? ms : 1) * sizeof(uint64_t));
  if (!ek) {
    orc_graph_free(g);
    return NULL;
  }
  int64_t k = 0;
  for (int64_t i = 0; i < ne; ++i) {
    if (v0[i] == v1[i]) continue;
    const uint64_t a = g->orig2new[v0[i]], b = g->orig2new[v1[i]];
    ek[k++] = (a << lg) | b;
    ek[k++] = (b << lg) | a;
  }
  radix_sort_u64(ek, ms, 2 * lg);
  int64_t m = 0;
  for (int64_t i = 0; i < ms; ++i)
    if (i == 0 || ek[i] != ek[i - 1]) ek[m++] = ek[i];
  g->m = m;
  g->col = (uint32_t*)malloc((size_t)(m ? m : 1) * sizeof(uint32_t));
  if (!g->col) {
    free(ek);
    orc_graph_free(g);
    return NULL;
  }
  const uint64_t lowmask = ((uint64_t)1 << lg) - 1;
  for (int64_t i = 0; i < m; ++i) {
    g->col[i] = (uint32_t)(ek[i] & lowmask);
    g->row_off[(ek[i] >> lg) + 1]++;
  }
  for (int64_t v = 0; v < n; ++v) g->row_off[v + 1] += g->row_off[v];
  free(ek);
  return g;
}

orc_graph* orc_graph_wrap(int log_n, int64_t n_c, int64_t m,
                          const int64_t* row_off, const uint32_t* col,
                          const int64_t* deg_new, const uint32_t* new2orig,
                          const uint32_t* orig2new) {
  orc_graph* g = (orc_graph*)calloc(1, sizeof(orc_graph));
  if (!g) return NULL;
  g->log_n = log_n;
  g->n = (int64_t)1 << log_n;
  g->n_c = n_c;
  g->m = m;
  g->m_sym = -1;
  g->row_off = (int64_t*)row_off;
  g->col = (uint32_t*)col;
  g->deg_new = (int64_t*)deg_new;
  g->new2orig = (uint32_t*)new2orig;
  g->orig2new = (uint32_t*)orig2new;
  g->max_used = 0;
  for (int64_t i = 0; i < n_c; ++i)
    if ((int64_t)new2orig[i] > g->max_used) g->max_used = new2orig[i];
  g->owns = 0;
  return g;
}

void orc_graph_free(orc_graph* g) {
  if (!g) return;
  if (g->owns) {
    free(g->row_off);
    free(g->col);
    free(g->deg_new);
    free(g->new2orig);
    free(g->orig2new);
  }
  free(g);
}

void orc_graph_get_info(const orc_graph* g, orc_graph_info* info) {
  info->log_n = g->log_n;
  info->n = g->n;
  info->n_c = g->n_c;
  info->m = g->m;
  info->m_sym = g->m_sym;
  info->max_used_vertex = g->max_used;
}

const int64_t* orc_graph_row_off(const orc_graph* g) { return g->row_off; }
const uint32_t* orc_graph_col(const orc_graph* g) { return g->col; }
const int64_t* orc_graph_deg_new(const orc_graph* g) { return g->deg_new; }
const uint32_t* orc_graph_new2orig(const orc_graph* g) { return g->new2orig; }
const uint32_t* orc_graph_orig2new(const orc_graph* g) { return g->orig2new; }

int orc_has_edge(const orc_graph* g, int64_t v) {
  if (v < 0 || v >= g->n) return 0;
  return g->orig2new[v] < (uint32_t)g->n_c;
}

/* ------------------------------------------------------------------------ */
/* BFS                                                                       */
/* ------------------------------------------------------------------------ */
#define PRED_ENTRY(parent, depth) \
  ((int64_t)(((uint64_t)(depth) << 48) | ((uint64_t)(parent)&0xFFFFFFFFFFFFull)))

int orc_bfs(const orc_graph* g, int64_t root, int32_t* level, int64_t* pred) {
  const int64_t n = g->n;
  for (int64_t v = 0; v < n; ++v) {
    level[v] = -1;
    if (pred) pred[v] = -1;
  }
  if (root < 0 || root >= n) return 0;
  uint32_t* q = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  int32_t* lev_new = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  for (int64_t v = 0; v < n; ++v) lev_new[v] = -1;
  const uint32_t r = g->orig2new[root];
  int64_t head = 0, tail = 0;
  q[tail++] = r;
  lev_new[r] = 0;
  if (pred) pred[root] = root;
  int maxlev = 0;
  while (head < tail) {
    const uint32_t u = q[head++];
    const int32_t lu = lev_new[u];
    for (int64_t e = g->row_off[u]; e < g->row_off[u + 1]; ++e) {
      const uint32_t v = g->col[e];
      if (lev_new[v] >= 0) continue;
      lev_new[v] = lu + 1;
      if (lu + 1 > maxlev) maxlev = lu + 1;
      if (pred) pred[g->new2orig[v]] = PRED_ENTRY(g->new2orig[u], lu + 1);
      q[tail++] = v;
    }
  }
  for (int64_t i = 0; i < n; ++i) level[g->new2orig[i]] = lev_new[i];
  free(q);
  free(lev_new);
  return maxlev + 1;
}

/* The decision half of the level loop, bfs.hpp:3298-3457.  Called after a
 * level finished; updates direction state and at_data. */
typedef struct {
  int next_td;  /* next_forward_or_backward_ */
  int growing;  /* growing_or_shrinking_ */
  int64_t prev; /* prev_global_nq_size */
} do_state;

static void do_decide(do_state* s, int td, int64_t nq_size, int64_t nq_edges,
                      int64_t m, int64_t n_c, int ef, double alpha,
                      double beta, int level, int64_t at[9]) {
  if (s->growing && nq_size > s->prev) {
    if (td) {
      if (nq_edges > (int64_t)(m / alpha)) {
        s->next_td = 0;
        at[0] = nq_edges; /* AUTO_GLOBAL_NQ_EDGES */
        at[1] = m;        /* AUTO_NUM_GLOBAL_EDGES */
        at[2] = level;    /* AUTO_T2B_LEVEL */
      } else {
        at[6] = nq_edges; /* AUTO_PRE_GLOBAL_NQ_EDGES */
      }
    }
  } else {
    if (!td) {
      if (nq_size < (int64_t)(n_c / (beta * ef * 2.0))) {
        s->next_td = 1;
        s->growing = 0;
        at[3] = nq_size; /* AUTO_GLOBAL_NQ_SIZE */
        at[4] = n_c;     /* AUTO_NUM_GLOBAL_VERTS */
        at[5] = level;   /* AUTO_B2T_LEVEL */
      } else {
        at[7] = nq_size; /* AUTO_PRE_GLOBAL_NQ_SIZE */
      }
    }
  }
}

int orc_do_schedule(int64_t m, int64_t n_c, int ef, double alpha, double beta,
                    int root_level, const int64_t* nq_size,
                    const int64_t* nq_edges, int nlev, int64_t at[9],
                    int8_t* dirs, int max_dirs) {
  do_state s = {1, 1, 0};
  int level = root_level, executed = 0;
  for (int i = 0; i < 9; ++i) at[i] = -1;
  for (int step = 1;; ++step) {
    at[8] = ++level;
    const int td = s.next_td;
    if (executed < max_dirs) dirs[executed] = (int8_t)td;
    ++executed;
    const int64_t ns = step < nlev ? nq_size[step] : 0;
    const int64_t ned = step < nlev ? nq_edges[step] : 0;
    if (ns == 0) break;
    do_decide(&s, td, ns, ned, m, n_c, ef, alpha, beta, level, at);
    s.prev = ns;
  }
  return executed;
}

int orc_do_bfs(const orc_graph* g, int64_t root, int ef, double alpha,
               double beta, int64_t* pred, int64_t at[9], int8_t* dirs,
               int max_dirs) {
  const int64_t n = g->n, nc = g->n_c;
  for (int64_t v = 0; v < n; ++v) pred[v] = -1;
  for (int i = 0; i < 9; ++i) at[i] = -1;
  if (root < 0 || root >= n) return 0;
  uint8_t* vis = (uint8_t*)calloc((size_t)(nc ? nc : 1), 1);
  uint8_t* vis_old = (uint8_t*)calloc((size_t)(nc ? nc : 1), 1);
  uint32_t* cq = (uint32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(uint32_t));
  uint32_t* nq = (uint32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(uint32_t));
  const uint32_t r = g->orig2new[root];
  int64_t cq_n = 0;
  pred[root] = root;
  if (r < (uint32_t)nc) {
    vis[r] = 1;
    cq[cq_n++] = r;
  }
  do_state s = {1, 1, 0};
  int level = 0, executed = 0;
  for (;;) {
    at[8] = ++level;
    const int td = s.next_td;
    if (executed < max_dirs) dirs[executed] = (int8_t)td;
    ++executed;
    int64_t nq_n = 0, nq_edges = 0;
    if (td) {
      for (int64_t i = 0; i < cq_n; ++i) {
        const uint32_t u = cq[i];
        for (int64_t e = g->row_off[u]; e < g->row_off[u + 1]; ++e) {
          const uint32_t v = g->col[e];
          if (vis[v]) continue;
          vis[v] = 1;
          pred[g->new2orig[v]] = PRED_ENTRY(g->new2orig[u], level);
          nq[nq_n++] = v;
          nq_edges += g->deg_new[v];
        }
      }
    } else {
      memcpy(vis_old, vis, (size_t)nc);
      for (int64_t v = 0; v < nc; ++v) {
        if (vis_old[v]) continue;
        for (int64_t e = g->row_off[v]; e < g->row_off[v + 1]; ++e) {
          const uint32_t u = g->col[e];
          if (!vis_old[u]) continue;
          vis[v] = 1;
          pred[g->new2orig[v]] = PRED_ENTRY(g->new2orig[u], level);
          nq[nq_n++] = (uint32_t)v;
          nq_edges += g->deg_new[v];
          break;
        }
      }
    }
    if (nq_n == 0) break;
    do_decide(&s, td, nq_n, nq_edges, g->m, nc, ef, alpha, beta, level, at);
    s.prev = nq_n;
    uint32_t* t = cq;
    cq = nq;
    nq = t;
    cq_n = nq_n;
  }
  free(vis);
  free(vis_old);
  free(cq);
  free(nq);
  return executed;
}

/* The same direction-optimizing BFS with OpenMP, as the reference's CPU path
 * runs it: top-down threads split the frontier and claim children
 * (top_down_search, bfs.hpp:1355-1395, per-thread queue buffers flushed to
 * the shared next queue), bottom-up threads split the vertices
 * (bottom_up_search_bitmap, bfs.hpp:2828-2978).  Frontier sets, and so
 * at_data, dirs and depths, do not depend on thread order and equal
 * orc_do_bfs's; a top-down parent is whichever frontier vertex claims the
 * child first (bottom-up parents are the first visited neighbour in row
 * order, as in orc_do_bfs). */
#define ORC_OMP_BUF 4096
int orc_do_bfs_omp(const orc_graph* g, int64_t root, int ef, double alpha,
                   double beta, int64_t* pred, int64_t at[9], int8_t* dirs,
                   int max_dirs, int nthreads) {
  const int64_t n = g->n, nc = g->n_c;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t v = 0; v < n; ++v) pred[v] = -1;
  for (int i = 0; i < 9; ++i) at[i] = -1;
  if (root < 0 || root >= n) return 0;
  uint8_t* vis = (uint8_t*)calloc((size_t)(nc ? nc : 1), 1);
  uint8_t* vis_old = (uint8_t*)calloc((size_t)(nc ? nc : 1), 1);
  uint32_t* cq = (uint32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(uint32_t));
  uint32_t* nq = (uint32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(uint32_t));
  const uint32_t r = g->orig2new[root];
  int64_t cq_n = 0;
  pred[root] = root;
  if (r < (uint32_t)nc) {
    vis[r] = 1;
    cq[cq_n++] = r;
  }
  do_state s = {1, 1, 0};
  int level = 0, executed = 0;
  for (;;) {
    at[8] = ++level;
    const int td = s.next_td;
    if (executed < max_dirs) dirs[executed] = (int8_t)td;
    ++executed;
    int64_t nq_n = 0, nq_edges = 0;
    if (!td) {
      const int64_t blk = 1 << 20;
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (int64_t b = 0; b < (nc + blk - 1) / blk; ++b) {
        const int64_t lo = b * blk, len = (nc - lo < blk) ? nc - lo : blk;
        memcpy(vis_old + lo, vis + lo, (size_t)len);
      }
    }
#pragma omp parallel num_threads(nthreads) reduction(+ : nq_edges)
    {
      uint32_t buf[ORC_OMP_BUF];
      int nb = 0;
#define ORC_FLUSH()                                                         \
  do {                                                                      \
    const int64_t at_ = __atomic_fetch_add(&nq_n, (int64_t)nb, __ATOMIC_RELAXED); \
    memcpy(nq + at_, buf, (size_t)nb * sizeof(uint32_t));                   \
    nb = 0;                                                                 \
  } while (0)
      if (td) {
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < cq_n; ++i) {
          const uint32_t u = cq[i];
          for (int64_t e = g->row_off[u]; e < g->row_off[u + 1]; ++e) {
            const uint32_t v = g->col[e];
            if (__atomic_load_n(&vis[v], __ATOMIC_RELAXED)) continue;
            if (__atomic_exchange_n(&vis[v], (uint8_t)1, __ATOMIC_RELAXED)) continue;
            pred[g->new2orig[v]] = PRED_ENTRY(g->new2orig[u], level);
            nq_edges += g->deg_new[v];
            buf[nb++] = v;
            if (nb == ORC_OMP_BUF) ORC_FLUSH();
          }
        }
      } else {
#pragma omp for schedule(dynamic, 4096)
        for (int64_t v = 0; v < nc; ++v) {
          if (vis_old[v]) continue;
          for (int64_t e = g->row_off[v]; e < g->row_off[v + 1]; ++e) {
            const uint32_t u = g->col[e];
            if (!vis_old[u]) continue;
            vis[v] = 1;
            pred[g->new2orig[v]] = PRED_ENTRY(g->new2orig[u], level);
            nq_edges += g->deg_new[v];
            buf[nb++] = (uint32_t)v;
            if (nb == ORC_OMP_BUF) ORC_FLUSH();
            break;
          }
        }
      }
      if (nb) ORC_FLUSH();
#undef ORC_FLUSH
    }
    if (nq_n == 0) break;
    do_decide(&s, td, nq_n, nq_edges, g->m, nc, ef, alpha, beta, level, at);
    s.prev = nq_n;
    uint32_t* t = cq;
    cq = nq;
    nq = t;
    cq_n = nq_n;
  }
  free(vis);
  free(vis_old);
  free(cq);
  free(nq);
  return executed;
}

/* ------------------------------------------------------------------------ */
/* Validation (validate.hpp:608-1255), single rank                           */
/* ------------------------------------------------------------------------ */
static int64_t pe_parent(int64_t e) { return (int64_t)((uint64_t)e << 16) >> 16; }
static uint32_t pe_depth(int64_t e) { return (uint32_t)(((uint64_t)e >> 48) & 0xFFFF); }
static void pe_set_depth(int64_t* e, uint32_t d) {
  *e = (int64_t)(((uint64_t)*e & 0xFFFFFFFFFFFFull) | ((uint64_t)(d & 0xFFFF) << 48));
}

int orc_validate(const int64_t* v0, const int64_t* v1, int64_t ne,
                 int64_t nglobal, int64_t nlocal, int corebfs, int64_t root,
                 int64_t* pred, int64_t* te_out, int64_t errs[ORC_ERR_NUM]) {
  for (int i = 0; i < ORC_ERR_NUM; ++i) errs[i] = 0;
  *te_out = 0;
  /* stage 1: value ranges (validate.hpp:995-1016) */
  for (int64_t i = 0; i < nlocal; ++i) {
    const int64_t p = pe_parent(pred[i]);
    if (p < -1 || p >= nglobal) errs[ORC_ERR_RANGE]++;
  }
  if (errs[ORC_ERR_RANGE]) return 0;
  if (root < 0 || root >= nlocal) {
    errs[ORC_ERR_ROOT]++;
    return 0;
  }
  /* stage 2: root is its own parent (validate.hpp:631-640) */
  if (pe_parent(pred[root]) != root) {
    errs[ORC_ERR_ROOT]++;
    return 0;
  }
  /* stage 3: nothing else is its own parent (validate.hpp:645-676) */
  for (int64_t i = 0; i < nlocal; ++i)
    if (i != root && pe_parent(pred[i]) != -1 && pe_parent(pred[i]) == i)
      errs[ORC_ERR_SELF_PARENT]++;
  if (errs[ORC_ERR_SELF_PARENT]) return 0;
  /* stage 4: depth map */
  if (corebfs) {
    /* validate.hpp:1040-1141: depth[v] = depth[parent]+1, iterated */
    pe_set_depth(&pred[root], 0);
    for (;;) {
      int changed = 0;
      for (int64_t i = 0; i < nlocal; ++i) {
        if (i == root) continue;
        const int64_t p = pe_parent(pred[i]);
        if (p < 0 || pe_depth(pred[i]) != 0xFFFF) continue;
        const uint32_t dp = pe_depth(pred[p]);
        if (dp == 0xFFFF) continue;
        pe_set_depth(&pred[i], dp + 1);
        changed = 1;
      }
      if (!changed) break;
    }
  } else {
    /* validate.hpp:1145-1255 */
    if (pe_depth(pred[root]) != 0) errs[ORC_ERR_ROOT]++;
    for (int64_t i = 0; i < nlocal; ++i) {
      const int64_t p = pe_parent(pred[i]);
      const uint32_t d = pe_depth(pred[i]);
      if (p == -1 && d != 0xFFFF) errs[ORC_ERR_DEPTH_MAP]++;
      else if (p != -1 && d == 0xFFFF) errs[ORC_ERR_DEPTH_MAP]++;
    }
    for (int64_t i = 0; i < nlocal; ++i) {
      if (i == root) continue;
      const int64_t p = pe_parent(pred[i]);
      if (p == -1) continue;
      const uint32_t dp = pe_depth(pred[p]);
      if (dp == 0xFFFF) errs[ORC_ERR_DEPTH_MAP]++;
      if (pe_depth(pred[i]) != dp + 1) errs[ORC_ERR_DEPTH_MAP]++;
    }
    if (errs[ORC_ERR_ROOT] || errs[ORC_ERR_DEPTH_MAP]) return 0;
  }
  /* stage 5: every input edge (validate.hpp:682-968) */
  uint8_t* pred_valid = (uint8_t*)calloc((size_t)(nlocal ? nlocal : 1), 1);
  int64_t te = 0;
  for (int64_t i = 0; i < ne; ++i) {
    const int64_t s = v0[i], t = v1[i];
    const int64_t ps = pred[s], pt = pred[t];
    const int ds = (int)pe_depth(ps), dt = (int)pe_depth(pt);
    if (ds != 0xFFFF && dt == 0xFFFF) errs[ORC_ERR_EDGE_TREE]++;
    else if (ds == 0xFFFF && dt != 0xFFFF) errs[ORC_ERR_EDGE_TREE]++;
    else if (ds - dt < -1 || ds - dt > 1) errs[ORC_ERR_EDGE_GAP]++;
    else if (ds != 0xFFFF) ++te;
    const int s_to_t = pe_parent(ps) == t, t_to_s = pe_parent(pt) == s;
    if ((s_to_t || t_to_s) && s != t &&
        ((s_to_t && dt + 1 != ds) || (t_to_s && ds + 1 != dt)))
      errs[ORC_ERR_PRED_DEPTH] += s_to_t + t_to_s;
    if (s_to_t) pred_valid[s] = 1;
    if (t_to_s) pred_valid[t] = 1;
  }
  for (int64_t i = 0; i < nlocal; ++i) {
    if (pe_parent(pred[i]) == -1) continue;
    if (!pred_valid[i] && i != root) errs[ORC_ERR_NO_PARENT_EDGE]++;
  }
  free(pred_valid);
  *te_out = te;
  for (int i = 0; i < ORC_ERR_NUM; ++i)
    if (errs[i]) return 0;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Roots and metric                                                          */
/* ------------------------------------------------------------------------ */
int orc_find_roots(const orc_graph* g, int nroots, int r1, int r2,
                   int64_t* roots) {
  const int64_t nglobal = (int64_t)1 << g->log_n;
  int64_t counter = 0;
  int idx;
  for (idx = 0; idx < nroots; ++idx) {
    int64_t root;
    for (;;) {
      double d[2];
      orc_make_random_numbers(2, (uint64_t)(2 + r1), (uint64_t)(3 + r2),
                              counter, d);
      root = (int64_t)((d[0] + d[1]) * (double)nglobal) % nglobal;
      counter += 2;
      if (counter > 2 * nglobal) break;
      int dup = 0;
      for (int i = 0; i < idx; ++i)
        if (roots[i] == root) {
          dup = 1;
          break;
        }
      if (dup) continue;
      if (orc_has_edge(g, root)) break;
    }
    roots[idx] = root;
  }
  return idx;
}

double orc_harmonic_mean_teps(const double* t, const double* te, int n) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += t[i] / te[i];
  return 1.0 / (s / n);
}

/* ------------------------------------------------------------------------ */
/* CoreBFS                                                                   */
/* ------------------------------------------------------------------------ */
int orc_corebfs_prune(const orc_graph* g, int64_t* tree_parent,
                      uint8_t* has_parent, uint8_t* in_core, int64_t stats[4]) {
  const int64_t n = g->n, nc = g->n_c;
  int64_t* deg = (int64_t*)malloc((size_t)(nc ? nc : 1) * sizeof(int64_t));
  uint8_t* alive = (uint8_t*)malloc((size_t)(nc ? nc : 1));
  int64_t* par_new = (int64_t*)malloc((size_t)(nc ? nc : 1) * sizeof(int64_t));
  int64_t* leaves = (int64_t*)malloc((size_t)(nc ? nc : 1) * sizeof(int64_t));
  int64_t* twig_t = (int64_t*)malloc((size_t)(nc ? nc : 1) * sizeof(int64_t));
  for (int64_t v = 0; v < nc; ++v) {
    deg[v] = g->row_off[v + 1] - g->row_off[v];
    alive[v] = deg[v] > 0;
    par_new[v] = -1;
  }
  int iters = 0;
  for (;;) {
    int64_t nl = 0;
    for (int64_t v = 0; v < nc; ++v)
      if (deg[v] == 1) leaves[nl++] = v;
    ++iters;
    if (nl == 0) break;
    /* twigs: first neighbour live at the iteration start */
    for (int64_t i = 0; i < nl; ++i) {
      const int64_t s = leaves[i];
      twig_t[i] = -1;
      for (int64_t e = g->row_off[s]; e < g->row_off[s + 1]; ++e)
        if (alive[g->col[e]]) {
          twig_t[i] = g->col[e];
          break;
        }
    }
    for (int64_t i = 0; i < nl; ++i) { /* update_by_source */
      if (twig_t[i] < 0) continue;
      const int64_t s = leaves[i];
      deg[s] = 0;
      par_new[s] = twig_t[i];
      alive[s] = 0;
    }
    for (int64_t i = 0; i < nl; ++i) { /* update_by_target */
      const int64_t t = twig_t[i];
      if (t < 0 || deg[t] == 0) continue;
      if (--deg[t] == 0) alive[t] = 0;
    }
  }
  int64_t ncore = 0, mcore = 0, ntree = 0;
  for (int64_t o = 0; o < n; ++o) {
    tree_parent[o] = -1;
    has_parent[o] = 0;
    in_core[o] = 0;
  }
  for (int64_t v = 0; v < nc; ++v) {
    const uint32_t o = g->new2orig[v];
    if (par_new[v] >= 0) {
      tree_parent[o] = g->new2orig[par_new[v]];
      has_parent[o] = 1;
      ++ntree;
    }
    if (alive[v]) {
      in_core[o] = 1;
      ++ncore;
      for (int64_t e = g->row_off[v]; e < g->row_off[v + 1]; ++e)
        if (alive[g->col[e]]) ++mcore;
    }
  }
  stats[0] = ncore;
  stats[1] = mcore;
  stats[2] = ntree;
  stats[3] = iters;
  free(deg);
  free(alive);
  free(par_new);
  free(leaves);
  free(twig_t);
  return 0;
}

int64_t orc_corebfs_reset_unreachable(const orc_graph* g, int64_t* tree_parent,
                                      const uint8_t* has_parent,
                                      const uint8_t* in_core, int64_t* hd_root) {
  const int64_t n = g->n, nc = g->n_c;
  /* highest core degree, ties by lowest original ID */
  int64_t best = -1, best_deg = -1;
  for (int64_t o = 0; o < n; ++o) {
    if (!in_core[o]) continue;
    const uint32_t v = g->orig2new[o];
    int64_t d = 0;
    for (int64_t e = g->row_off[v]; e < g->row_off[v + 1]; ++e)
      if (in_core[g->new2orig[g->col[e]]]) ++d;
    if (d > best_deg) {
      best_deg = d;
      best = o;
    }
  }
  *hd_root = best;
  uint8_t* reach = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
  if (best >= 0) { /* BFS inside the core */
    uint32_t* q = (uint32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(uint32_t));
    int64_t h = 0, t = 0;
    q[t++] = g->orig2new[best];
    reach[best] = 1;
    while (h < t) {
      const uint32_t u = q[h++];
      for (int64_t e = g->row_off[u]; e < g->row_off[u + 1]; ++e) {
        const uint32_t w = g->col[e];
        const uint32_t ow = g->new2orig[w];
        if (!in_core[ow] || reach[ow]) continue;
        reach[ow] = 1;
        q[t++] = w;
      }
    }
    free(q);
  }
  /* reachability flows from parents to children through the forest */
  for (int changed = 1; changed;) {
    changed = 0;
    for (int64_t o = 0; o < n; ++o)
      if (has_parent[o] && !reach[o] && tree_parent[o] >= 0 && reach[tree_parent[o]]) {
        reach[o] = 1;
        changed = 1;
      }
  }
  int64_t removed = 0;
  for (int64_t o = 0; o < n; ++o)
    if (has_parent[o] && !reach[o]) {
      tree_parent[o] = -1;
      ++removed;
    }
  free(reach);
  return removed;
}

int orc_corebfs_run(const orc_graph* g, const int64_t* tree_parent,
                    const uint8_t* has_parent, const uint8_t* in_core,
                    int64_t root, int64_t* pred) {
  const int64_t n = g->n, nc = g->n_c;
  for (int64_t o = 0; o < n; ++o) pred[o] = -1;
  /* walk to the core: (vertex, entry) pairs along the path */
  int64_t pv[256], pe[256];
  int np = 0;
  int64_t u = root, p = root;
  int lv = 0;
  for (;; ++lv) {
    if (np < 256) {
      pv[np] = p;
      pe[np] = PRED_ENTRY(u, lv);
      ++np;
    }
    u = p;
    p = has_parent[u] ? tree_parent[u] : -1;
    if (p < 0) break;
  }
  const int64_t core_root = u;
  /* BFS inside the core from core_root at level lv (core vertices only) */
  if (in_core[core_root]) {
    uint32_t* q = (uint32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(uint32_t));
    int32_t* lev = (int32_t*)malloc((size_t)(nc ? nc : 1) * sizeof(int32_t));
    for (int64_t v = 0; v < nc; ++v) lev[v] = -1;
    int64_t h = 0, t = 0;
    const uint32_t r = g->orig2new[core_root];
    q[t++] = r;
    lev[r] = lv;
    pred[core_root] = core_root;
    while (h < t) {
      const uint32_t x = q[h++];
      for (int64_t e = g->row_off[x]; e < g->row_off[x + 1]; ++e) {
        const uint32_t w = g->col[e];
        if (!in_core[g->new2orig[w]] || lev[w] >= 0) continue;
        lev[w] = lev[x] + 1;
        pred[g->new2orig[w]] = PRED_ENTRY(g->new2orig[x], lev[w]);
        q[t++] = w;
      }
    }
    free(q);
    free(lev);
  }
  /* tree parents with unknown depth, then the walked path */
  for (int64_t o = 0; o < n; ++o)
    if (has_parent[o]) pred[o] = tree_parent[o] < 0 ? -1 : PRED_ENTRY(tree_parent[o], 0xFFFF);
  for (int i = 0; i < np; ++i) pred[pv[i]] = pe[i];
  return lv;
}
