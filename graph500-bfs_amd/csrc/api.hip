// api.hip -- the extern "C" entry points declared in include/g500bfs.h.
//
// Each entry point sets the handle's device, runs the C++ implementation and
// turns exceptions into status codes plus a per-thread message.
#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <new>
#include <string>

#include "engine.hpp"
#include "mrg.hpp"
#include "msbfs.hpp"

namespace g500 {

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

}  // namespace g500

namespace {

thread_local std::string t_last_error;

template <class F>
int guard(F&& f) {
  // launch checks read hipGetLastError: drop any error an earlier call left
  // behind (e.g. a best-effort call in a destructor) so that it is not
  // reported against this entry point
  (void)hipGetLastError();
  try {
    f();
    return G500_OK;
  } catch (const g500::Error& e) {
    t_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    t_last_error = "host allocation failed";
    return G500_ERR_NOMEM;
  } catch (const std::exception& e) {
    t_last_error = e.what();
    return G500_ERR_STATE;
  }
}

void use_device(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    g500::fail(G500_ERR_NODEV, "no HIP device available");
  if (device < 0 || device >= count)
    g500::fail(G500_ERR_ARG, "device %d outside [0, %d)", device, count);
  G500_CHECK_HIP(hipSetDevice(device));
}

g500_graph* new_graph(int device) {
  use_device(device);
  auto* g = new g500_graph;
  g->device = device;
  if (hipStreamCreate(&g->stream) != hipSuccess) {
    delete g;
    g500::fail(G500_ERR_HIP, "hipStreamCreate failed");
  }
  return g;
}

void need_graph(const g500_graph* g, bool constructed) {
  if (!g) g500::fail(G500_ERR_ARG, "null graph");
  if (constructed && !g->constructed)
    g500::fail(G500_ERR_STATE, "graph not constructed (call g500_graph_construct)");
}

void need_msbfs(g500_msbfs* b) {
  if (!b) g500::fail(G500_ERR_ARG, "null batched-BFS handle");
  G500_CHECK_HIP(hipSetDevice(b->g->device));
}
void need_result(g500_msbfs* b) {
  need_msbfs(b);
  if (b->in_flight) g500::fail(G500_ERR_STATE, "batch still in flight (call g500_msbfs_wait)");
  if (!b->have_result) g500::fail(G500_ERR_STATE, "no completed batch");
}
bool has_edge_host(g500_graph* g, int64_t v) {
  if (v < 0 || v >= g->n) return false;
  g500::ensure_host_orig2new(g);
  if ((int64_t)g->h_orig2new[(size_t)v] < g->n_c) return true;
  return g500::has_parent_host(g, v);
}

// Edges [0, count) of (v0, v1) to the device after the graph's current ones;
// every endpoint must fit 32 bits.  Copies through a bounded host staging
// buffer, so a caller's chunk of any size costs at most 8 Mi packed edges of
// host memory.
void append_edges(g500_graph* g, const int64_t* v0, const int64_t* v1, int64_t count) {
  if (count < 0 || (count > 0 && (!v0 || !v1))) g500::fail(G500_ERR_ARG, "bad edge arrays");
  if (g->constructed) g500::fail(G500_ERR_STATE, "graph already constructed");
  if (count > g->capacity - g->nedges)
    g500::fail(G500_ERR_ARG, "%lld more edges exceed the capacity of %lld (%lld held)",
               (long long)count, (long long)g->capacity, (long long)g->nedges);
  constexpr int64_t kStage = int64_t(1) << 23;
  std::vector<uint2> packed((size_t)std::min(count, kStage));
  for (int64_t b0 = 0; b0 < count; b0 += kStage) {
    const int64_t nb = std::min(kStage, count - b0);
    uint64_t ors = 0;
    for (int64_t i = 0; i < nb; ++i) {
      const int64_t a = v0[b0 + i], c = v1[b0 + i];
      if (a < 0 || c < 0 || a > 0xFFFFFFFFll || c > 0xFFFFFFFFll)
        g500::fail(G500_ERR_ARG, "edge %lld has an endpoint outside [0, 2^32)",
                   (long long)(g->nedges + i));
      packed[(size_t)i] = make_uint2((uint32_t)a, (uint32_t)c);
      ors |= (uint64_t)(a | c);
    }
    G500_CHECK_HIP(hipMemcpy(g->edges.get() + g->nedges, packed.data(), nb * sizeof(uint2),
                             hipMemcpyHostToDevice));
    g->nedges += nb;
    g->id_or |= ors;
  }
  g->id_bits = g->id_or ? 64 - __builtin_clzll(g->id_or) : 0;
}

g500_graph* new_graph_with_capacity(int device, int64_t capacity) {
  if (capacity < 0) g500::fail(G500_ERR_ARG, "negative edge capacity");
  g500_graph* g = new_graph(device);
  try {
    g->edges.alloc((size_t)capacity);
    g->capacity = capacity;
  } catch (...) {
    delete g;
    throw;
  }
  return g;
}

}  // namespace

g500_graph::~g500_graph() {
  if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

int g500_graph_generate(int device, int scale, int edgefactor, uint64_t seed1,
                        uint64_t seed2, g500_graph** out) {
  return guard([&] {
    if (!out) g500::fail(G500_ERR_ARG, "null output pointer");
    *out = nullptr;
    g500_graph* g = new_graph(device);
    try {
      g500::generate_edges(g, scale, edgefactor, seed1, seed2);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int g500_graph_create(int device, int64_t capacity, g500_graph** out) {
  return guard([&] {
    if (!out) g500::fail(G500_ERR_ARG, "null output pointer");
    *out = nullptr;
    *out = new_graph_with_capacity(device, capacity);
  });
}

int g500_graph_append_edges(g500_graph* g, const int64_t* v0, const int64_t* v1,
                            int64_t count) {
  return guard([&] {
    need_graph(g, false);
    use_device(g->device);
    append_edges(g, v0, v1, count);
  });
}

int g500_graph_from_edges(int device, const int64_t* v0, const int64_t* v1,
                          int64_t nedges, g500_graph** out) {
  return guard([&] {
    if (!out) g500::fail(G500_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (nedges < 0 || (nedges > 0 && (!v0 || !v1)))
      g500::fail(G500_ERR_ARG, "bad edge arrays");
    g500_graph* g = new_graph_with_capacity(device, nedges);
    try {
      append_edges(g, v0, v1, nedges);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int g500_graph_construct(g500_graph* g, int flags) {
  return guard([&] {
    need_graph(g, false);
    use_device(g->device);
    if (g->constructed) g500::fail(G500_ERR_STATE, "graph already constructed");
    if (flags & ~(G500_CONSTRUCT_COREBFS | G500_CONSTRUCT_SYMMETRIC_INPUT))
      g500::fail(G500_ERR_ARG, "unknown construct flags 0x%x", flags);
    g->symmetric_input = (flags & G500_CONSTRUCT_SYMMETRIC_INPUT) != 0;
    if (flags & G500_CONSTRUCT_COREBFS) {
      g500::corebfs_prune(g);
    } else {
      g500::construct_graph(g, 0);
    }
  });
}

int g500_graph_get_info(const g500_graph* g, g500_graph_info* info) {
  return guard([&] {
    need_graph(g, false);
    if (!info) g500::fail(G500_ERR_ARG, "null info");
    std::memset(info, 0, sizeof(*info));
    info->log_n = g->log_n;
    info->corebfs = g->corebfs;
    info->n = g->n;
    info->n_c = g->n_c;
    info->m = g->m;
    info->m_sym = g->m_sym;
    info->nedges_input = g->nedges;
    info->max_used_vertex = g->max_used;
    info->max_degree = g->max_deg;
    info->core_vertices = g->core_vertices;
    info->tree_vertices = g->tree_vertices;
    info->device_bytes = g->device_bytes();
    info->removed_parents = g->removed_parents;
    info->hd_root = g->corebfs ? g->hd_root : -1;
  });
}

int g500_graph_copy_forest(const g500_graph* g, int64_t* tree_parent, uint8_t* has_parent) {
  return guard([&] {
    need_graph(g, true);
    if (!g->corebfs) g500::fail(G500_ERR_STATE, "graph was not built with CoreBFS");
    use_device(g->device);
    std::vector<uint32_t> tp((size_t)g->n);
    G500_CHECK_HIP(hipMemcpy(tp.data(), g->tree_parent.get(), g->n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
    for (int64_t o = 0; o < g->n; ++o) {
      if (tree_parent) tree_parent[o] = tp[(size_t)o] == g500::kNoParent ? -1 : (int64_t)tp[(size_t)o];
      if (has_parent) has_parent[o] = g500::has_parent_host(g, o) ? 1 : 0;
    }
  });
}

int g500_graph_has_edge(g500_graph* g, int64_t v, int* out) {
  return guard([&] {
    need_graph(g, true);
    if (!out) g500::fail(G500_ERR_ARG, "null output");
    use_device(g->device);
    *out = has_edge_host(g, v) ? 1 : 0;
  });
}

int g500_graph_copy_edges(const g500_graph* g, int64_t start, int64_t count,
                          int64_t* v0, int64_t* v1) {
  return guard([&] {
    need_graph(g, false);
    if (start < 0 || count < 0 || start + count > g->nedges)
      g500::fail(G500_ERR_ARG, "edge range outside [0, %lld)", (long long)g->nedges);
    if (count && (!v0 || !v1)) g500::fail(G500_ERR_ARG, "null output arrays");
    use_device(g->device);
    std::vector<uint2> tmp((size_t)count);
    if (count)
      G500_CHECK_HIP(hipMemcpy(tmp.data(), g->edges.get() + start, count * sizeof(uint2),
                               hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < count; ++i) {
      v0[i] = tmp[(size_t)i].x;
      v1[i] = tmp[(size_t)i].y;
    }
  });
}

int g500_graph_copy_csr(const g500_graph* g, int64_t* row_off, uint32_t* col,
                        int64_t* deg_new, uint32_t* new2orig, uint32_t* orig2new) {
  return guard([&] {
    need_graph(g, true);
    use_device(g->device);
    if (row_off) {
      G500_CHECK_HIP(hipMemcpy(row_off, g->row_off.get(), (g->n_c + 1) * sizeof(int64_t),
                               hipMemcpyDeviceToHost));
      for (int64_t v = g->n_c + 1; v <= g->n; ++v) row_off[v] = g->m;
    }
    if (col && g->m)
      G500_CHECK_HIP(hipMemcpy(col, g->col.get(), g->m * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
    if (deg_new) {
      std::vector<uint32_t> d((size_t)g->n);
      G500_CHECK_HIP(hipMemcpy(d.data(), g->deg_new.get(), g->n * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < g->n; ++i) deg_new[i] = d[(size_t)i];
    }
    if (new2orig)
      G500_CHECK_HIP(hipMemcpy(new2orig, g->new2orig.get(), g->n * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
    if (orig2new)
      G500_CHECK_HIP(hipMemcpy(orig2new, g->orig2new.get(), g->n * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
  });
}

// mpi/benchmark_helper.hpp:527-562 at one rank.
int g500_find_roots(g500_graph* g, int nroots, int r1, int r2, int64_t* roots,
                    int* found) {
  return guard([&] {
    need_graph(g, true);
    if (nroots < 0 || (nroots && !roots) || !found) g500::fail(G500_ERR_ARG, "bad arguments");
    use_device(g->device);
    const int64_t nglobal = (int64_t)1 << g->log_n;
    int64_t counter = 0;
    int idx;
    for (idx = 0; idx < nroots; ++idx) {
      int64_t root;
      for (;;) {
        double d[2];
        g500::make_random_numbers_host(2, (uint64_t)(2 + r1), (uint64_t)(3 + r2), counter, d);
        root = (int64_t)((d[0] + d[1]) * (double)nglobal) % nglobal;
        counter += 2;
        if (counter > 2 * nglobal) break;
        bool dup = false;
        for (int i = 0; i < idx; ++i)
          if (roots[i] == root) {
            dup = true;
            break;
          }
        if (dup) continue;
        if (has_edge_host(g, root)) break;
      }
      roots[idx] = root;
    }
    *found = idx;
  });
}

int g500_make_random_numbers(int64_t nvalues, uint64_t seed1, uint64_t seed2,
                             int64_t position, double* out) {
  return guard([&] {
    if (nvalues < 0 || (nvalues && !out) || position < 0)
      g500::fail(G500_ERR_ARG, "bad arguments");
    g500::make_random_numbers_host(nvalues, seed1, seed2, position, out);
  });
}

int g500_graph_free(g500_graph* g) {
  return guard([&] {
    if (!g) return;
    // an engine keeps pointers into the graph (and reads g->device when it
    // is destroyed): freeing the graph first would leave them dangling
    if (g->live_engines > 0)
      g500::fail(G500_ERR_STATE, "graph still has %d engine(s): destroy them first",
                 g->live_engines);
    (void)hipSetDevice(g->device);
    (void)hipStreamSynchronize(g->stream);
    delete g;
  });
}

int g500_bfs_create(g500_graph* g, g500_bfs** out) {
  return g500_bfs_create_ex(g, 16, 10.0, 14.0, out);
}

int g500_bfs_create_ex(g500_graph* g, int edgefactor, double alpha, double beta,
                       g500_bfs** out) {
  return guard([&] {
    need_graph(g, true);
    if (!out) g500::fail(G500_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!(alpha > 0) || !(beta > 0) || edgefactor <= 0)
      g500::fail(G500_ERR_ARG, "alpha, beta and edgefactor must be positive");
    use_device(g->device);
    auto* b = new g500_bfs;
    try {
      b->g = g;
      G500_CHECK_HIP(hipStreamCreate(&b->stream));
      g500::bfs_init(b, edgefactor, alpha, beta);
    } catch (...) {
      delete b;
      throw;
    }
    ++g->live_engines;
    *out = b;
  });
}

int g500_bfs_run(g500_bfs* b, int64_t root, int edgefactor, double alpha,
                 double beta, int64_t* at_data) {
  return guard([&] {
    if (!b) g500::fail(G500_ERR_ARG, "null bfs handle");
    use_device(b->g->device);
    g500::bfs_run(b, root, edgefactor, alpha, beta, at_data);
  });
}

int g500_bfs_get_pred(g500_bfs* b, int64_t* pred) {
  return guard([&] {
    if (!b || !pred) g500::fail(G500_ERR_ARG, "null argument");
    if (!b->have_result) g500::fail(G500_ERR_STATE, "no BFS result");
    use_device(b->g->device);
    G500_CHECK_HIP(hipMemcpyAsync(pred, b->pred.get(), b->g->n * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, b->stream));
    G500_CHECK_HIP(hipStreamSynchronize(b->stream));
  });
}

int g500_bfs_get_levels(g500_bfs* b, uint16_t* levels) {
  return guard([&] {
    if (!b || !levels) g500::fail(G500_ERR_ARG, "null argument");
    if (!b->have_result) g500::fail(G500_ERR_STATE, "no BFS result");
    use_device(b->g->device);
    g500::bfs_levels(b, levels);
  });
}

int g500_bfs_validate(g500_bfs* b, int64_t root, const int64_t* host_pred,
                      int64_t* te, int64_t* errs, int* ok) {
  return guard([&] {
    if (!b || !te || !errs || !ok) g500::fail(G500_ERR_ARG, "null argument");
    use_device(b->g->device);
    g500::bfs_validate(b, root, host_pred, te, errs, ok);
  });
}

int g500_bfs_set_timing(g500_bfs* b, int on) {
  return guard([&] {
    if (!b) g500::fail(G500_ERR_ARG, "null bfs handle");
    b->timing = on != 0;
  });
}

int g500_bfs_set_exec_alpha(g500_bfs* b, double exec_alpha) {
  return guard([&] {
    if (!b) g500::fail(G500_ERR_ARG, "null bfs handle");
    if (!(exec_alpha >= 0)) g500::fail(G500_ERR_ARG, "exec_alpha must be >= 0");
    b->exec_alpha = exec_alpha;
  });
}

int g500_bfs_set_counting(g500_bfs* b, int on) {
  return guard([&] {
    if (!b) g500::fail(G500_ERR_ARG, "null bfs handle");
    b->counting = on != 0;
  });
}

int g500_bfs_get_stats(g500_bfs* b, g500_run_stats* st, g500_level_trace* levels,
                       int max_levels) {
  return guard([&] {
    if (!b) g500::fail(G500_ERR_ARG, "null bfs handle");
    if (st) *st = b->stats;
    if (levels)
      for (int i = 0; i < max_levels && i < (int)b->trace.size(); ++i)
        levels[i] = b->trace[(size_t)i];
  });
}

int g500_bfs_destroy(g500_bfs* b) {
  return guard([&] {
    if (!b) return;
    (void)hipSetDevice(b->g->device);
    // a run that failed inside a level can leave kernels queued on either
    // stream (the pred gather runs on gstream): drain both before the
    // buffers and the pinned report they write are released
    (void)hipStreamSynchronize(b->stream);
    if (b->gstream) (void)hipStreamSynchronize(b->gstream);
    --b->g->live_engines;
    delete b;
  });
}

int g500_msbfs_footprint(const g500_graph* g, int max_roots, int flags, int64_t* bytes) {
  return guard([&] {
    need_graph(g, true);
    if (!bytes) g500::fail(G500_ERR_ARG, "null output pointer");
    if (max_roots < 1 || max_roots > G500_MSBFS_MAX_ROOTS)
      g500::fail(G500_ERR_ARG, "max_roots %d outside [1, %d]", max_roots, G500_MSBFS_MAX_ROOTS);
    if (flags & ~G500_MSBFS_NO_LEVEL_ARRAYS) g500::fail(G500_ERR_ARG, "unknown flags 0x%x", flags);
    *bytes = g500::msbfs_footprint(g, max_roots, !(flags & G500_MSBFS_NO_LEVEL_ARRAYS));
  });
}

int g500_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
  return guard([&] {
    if (!free_bytes || !total_bytes) g500::fail(G500_ERR_ARG, "null output pointer");
    use_device(device);
    size_t f = 0, t = 0;
    G500_CHECK_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
  });
}

int g500_msbfs_create(g500_graph* g, int max_roots, g500_msbfs** out) {
  return g500_msbfs_create_ex(g, max_roots, 0, out);
}

int g500_msbfs_create_ex(g500_graph* g, int max_roots, int flags, g500_msbfs** out) {
  return guard([&] {
    if (flags & ~G500_MSBFS_NO_LEVEL_ARRAYS) g500::fail(G500_ERR_ARG, "unknown flags 0x%x", flags);
    need_graph(g, true);
    if (!out) g500::fail(G500_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (max_roots < 1 || max_roots > G500_MSBFS_MAX_ROOTS)
      g500::fail(G500_ERR_ARG, "max_roots %d outside [1, %d]", max_roots, G500_MSBFS_MAX_ROOTS);
    if (g->corebfs)
      g500::fail(G500_ERR_STATE, "batched BFS needs a graph constructed without CoreBFS");
    use_device(g->device);
    auto* b = new g500_msbfs;
    try {
      b->g = g;
      g500::msbfs_init(b, max_roots, !(flags & G500_MSBFS_NO_LEVEL_ARRAYS));
    } catch (...) {
      delete b;
      throw;
    }
    ++g->live_engines;
    *out = b;
  });
}

int g500_msbfs_set_tuning(g500_msbfs* b, const g500_msbfs_tuning* t) {
  return guard([&] {
    need_msbfs(b);
    if (!t) g500::fail(G500_ERR_ARG, "null tuning");
    if (b->in_flight) g500::fail(G500_ERR_STATE, "batch in flight");
    g500::msbfs_set_tuning(b, *t);
  });
}

int g500_msbfs_get_tuning(g500_msbfs* b, g500_msbfs_tuning* t) {
  return guard([&] {
    need_msbfs(b);
    if (!t) g500::fail(G500_ERR_ARG, "null tuning");
    *t = b->tuning;
  });
}

int g500_msbfs_launch(g500_msbfs* b, const int64_t* roots, int nroots) {
  return guard([&] {
    need_msbfs(b);
    g500::msbfs_launch(b, roots, nroots);
  });
}

int g500_msbfs_query(g500_msbfs* b, int* done) {
  return guard([&] {
    need_msbfs(b);
    if (!done) g500::fail(G500_ERR_ARG, "null output pointer");
    *done = g500::msbfs_query(b) ? 1 : 0;
  });
}

int g500_msbfs_wait(g500_msbfs* b) {
  return guard([&] {
    need_msbfs(b);
    g500::msbfs_wait(b);
  });
}

int g500_msbfs_run(g500_msbfs* b, const int64_t* roots, int nroots) {
  return guard([&] {
    need_msbfs(b);
    g500::msbfs_launch(b, roots, nroots);
    g500::msbfs_wait(b);
  });
}

int g500_msbfs_get_levels(g500_msbfs* b, int index, uint8_t* host_levels) {
  return guard([&] {
    need_result(b);
    if (!host_levels) g500::fail(G500_ERR_ARG, "null output buffer");
    if (!b->level_arrays)
      g500::fail(G500_ERR_STATE, "engine created without level arrays (use point lookups)");
    if (index < 0 || index >= b->nroots)
      g500::fail(G500_ERR_ARG, "root index %d outside [0, %d)", index, b->nroots);
    const size_t n = (size_t)b->g->n;
    G500_CHECK_HIP(hipMemcpy(host_levels, b->out.get() + (size_t)index * n, n,
                             hipMemcpyDeviceToHost));
  });
}

int g500_msbfs_device_levels(g500_msbfs* b, void** dev_ptr, int64_t* row_stride) {
  return guard([&] {
    need_result(b);
    if (!dev_ptr || !row_stride) g500::fail(G500_ERR_ARG, "null output pointer");
    if (!b->level_arrays)
      g500::fail(G500_ERR_STATE, "engine created without level arrays (use point lookups)");
    *dev_ptr = b->out.get();
    *row_stride = b->g->n;
  });
}

int g500_msbfs_gather_levels(g500_msbfs* b, const int64_t* pairs, int64_t count,
                             uint8_t* levels) {
  return guard([&] {
    need_result(b);
    if (count < 0 || (count > 0 && (!pairs || !levels)))
      g500::fail(G500_ERR_ARG, "bad pair list");
    g500::msbfs_gather(b, pairs, count, levels);
  });
}

int g500_msbfs_gather_levels_noalloc(g500_msbfs* b, const int64_t* pairs, int64_t count,
                                     uint8_t* levels) {
  return guard([&] {
    need_result(b);
    if (count < 0 || (count > 0 && (!pairs || !levels)))
      g500::fail(G500_ERR_ARG, "bad pair list");
    g500::msbfs_gather_noalloc(b, pairs, count, levels);
  });
}

int g500_msbfs_autotune(g500_msbfs* b, const int64_t* roots, int nroots, int rounds,
                        g500_msbfs_tuning* best, double* best_ms) {
  return guard([&] {
    need_msbfs(b);
    if (b->in_flight) g500::fail(G500_ERR_STATE, "batch in flight");
    if (rounds < 1 || rounds > 8) g500::fail(G500_ERR_ARG, "rounds %d outside [1, 8]", rounds);
    const double ms = g500::msbfs_autotune(b, roots, nroots, rounds, best);
    if (best_ms) *best_ms = ms;
  });
}

int g500_msbfs_get_te(g500_msbfs* b, int64_t* te) {
  return guard([&] {
    need_result(b);
    if (!te) g500::fail(G500_ERR_ARG, "null output buffer");
    G500_CHECK_HIP(hipMemcpy(te, b->te.get(), (size_t)b->nroots * sizeof(int64_t),
                             hipMemcpyDeviceToHost));
  });
}

int g500_msbfs_get_stats(g500_msbfs* b, g500_msbfs_stats* st) {
  return guard([&] {
    need_result(b);
    if (!st) g500::fail(G500_ERR_ARG, "null output pointer");
    *st = b->stats;
  });
}

int g500_msbfs_destroy(g500_msbfs* b) {
  return guard([&] {
    if (!b) return;
    (void)hipSetDevice(b->g->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    --b->g->live_engines;
    delete b;
  });
}

const char* g500_last_error(void) { return t_last_error.c_str(); }

}  // extern "C"
