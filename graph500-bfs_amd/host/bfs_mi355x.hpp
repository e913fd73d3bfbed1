// bfs_mi355x.hpp -- BfsOnMI355X: the reference's BfsBase method set
// (mpi/bfs.hpp:116-172) over the C ABI of libg500bfs_hip.so.
//
// Drop-in use in the reference driver (mpi/main.cc):
//   typedef BfsOnMI355X BfsOnCPU;   // instead of mpi/bfs_cpu.hpp:13
// construct / prepare_bfs / run_bfs / get_pred / end_bfs / has_edge and the
// public graph_ fields keep their names, arguments and meaning.  pred stays
// on the GPU during run_bfs; get_pred copies it out (outside the timer, as
// mpi/main.cc:626-641 expects).  Errors are thrown as g500host::Error
// (the reference throws; there are no status codes at this level).
#pragma once

#include <g500bfs.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace g500host {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char* what) {
  if (rc != G500_OK)
    throw Error(rc, std::string(what) + ": " + g500_last_error());
}

// The Graph2DCSR fields the driver reads (mpi/graph_constructor.hpp:115-333).
struct GraphView {
  g500_graph* h = nullptr;
  int log_orig_global_verts_ = 0;  // estimated SCALE
  int64_t num_orig_local_verts_ = 0;
  int64_t num_global_edges_ = 0;   // deduplicated directed CSR edges
  int64_t num_global_verts_ = 0;   // vertices with at least one edge
  int64_t max_used_vertex_ = 0;
  int reorder_bits_ = 0;           // printed under SMALL_REORDER_BIT (mpi/main.cc:691-694)

  int64_t pred_size() const { return num_orig_local_verts_; }
  int log_orig_global_verts() const { return log_orig_global_verts_; }
  bool has_edge(int64_t v) const {
    int out = 0;
    check(g500_graph_has_edge(h, v, &out), "has_edge");
    return out != 0;
  }
};

class BfsOnMI355X {
 public:
  explicit BfsOnMI355X(int device = 0) : device_(device) {}
  ~BfsOnMI355X() {
    end_bfs();
    if (graph_.h) g500_graph_free(graph_.h);
  }
  BfsOnMI355X(const BfsOnMI355X&) = delete;
  BfsOnMI355X& operator=(const BfsOnMI355X&) = delete;

  // BfsBase::construct(scale, corebfs, sym_edge_list) (mpi/bfs.hpp:116-139).
  // EdgeList follows the reference's EdgeListStorage read protocol
  // (beginRead / read / endRead, generator/graph_generator.hpp:46-485) and
  // holds the output of make_symmetry_edge_list.
  // Each chunk goes straight to device memory sized from num_local_edges();
  // the host holds one chunk's endpoints at a time.
  template <typename EdgeList>
  void construct(int /*scale*/, bool corebfs, EdgeList* sym_edge_list) {
    reset_graph();
    check(g500_graph_create(device_, sym_edge_list->num_local_edges(), &graph_.h),
          "graph create");
    std::vector<int64_t> v0, v1;
    const int chunks = sym_edge_list->beginRead(false);
    for (int c = 0; c < chunks; ++c) {
      typename EdgeList::edge_type* buf = nullptr;
      const int n = sym_edge_list->read(&buf);
      v0.resize((size_t)n);
      v1.resize((size_t)n);
      for (int i = 0; i < n; ++i) {
        v0[(size_t)i] = buf[i].v0();
        v1[(size_t)i] = buf[i].v1();
      }
      check(g500_graph_append_edges(graph_.h, v0.data(), v1.data(), n), "graph upload");
    }
    sym_edge_list->endRead();
    finish_construct(corebfs, true);
  }

  // Edge list already in memory (symmetric or not).
  void construct_from_edges(const int64_t* v0, const int64_t* v1, int64_t n,
                            bool corebfs, bool symmetric) {
    reset_graph();
    check(g500_graph_from_edges(device_, v0, v1, n, &graph_.h), "graph upload");
    finish_construct(corebfs, symmetric);
  }

  // Graph500 Kronecker graph generated on the GPU (the 13-26 GB edge list
  // never crosses PCIe); separate steps so the driver can time them apart.
  void generate(int scale, int edgefactor, uint64_t seed1, uint64_t seed2) {
    reset_graph();
    check(g500_graph_generate(device_, scale, edgefactor, seed1, seed2, &graph_.h),
          "generate");
  }
  void construct_generated(bool corebfs) { finish_construct(corebfs, false); }

  // BfsBase::prepare_bfs (mpi/bfs.hpp:141-151).
  void prepare_bfs(int validation_level, bool pre_exec, bool real_benchmark,
                   int edgefactor, double alpha, double beta, int64_t* /*pred*/) {
    printInformation(validation_level, pre_exec, real_benchmark);
    end_bfs();
    // with CoreBFS the reachability BFS runs with these settings, as
    // reset_unreachable does (mpi/corebfs_adaptor.hpp:228-249)
    check(g500_bfs_create_ex(graph_.h, edgefactor, alpha, beta, &engine_), "prepare_bfs");
    if (getenv("G500_PROFILE_LEVELS")) check(g500_bfs_set_timing(engine_, 1), "timing");
  }

  // BfsBase::run_bfs (mpi/bfs.hpp:153-155, 3214-3232).
  void run_bfs(int64_t root, int64_t* /*pred*/, int edgefactor, double alpha,
               double beta, int64_t* auto_tuning_data) {
    check(g500_bfs_run(engine_, root, edgefactor, alpha, beta, auto_tuning_data),
          "run_bfs");
  }

  // BfsBase::get_pred (mpi/bfs.hpp:161-163): device pred -> host.
  void get_pred(int64_t* pred) { check(g500_bfs_get_pred(engine_, pred), "get_pred"); }

  // BfsBase::end_bfs (mpi/bfs.hpp:165).
  void end_bfs() {
    if (engine_) g500_bfs_destroy(engine_);
    engine_ = nullptr;
  }

  // BfsBase::has_edge (mpi/bfs.hpp:167-170).
  bool has_edge(int64_t v) const { return graph_.has_edge(v); }

  // validate_bfs_result (mpi/validate.hpp:1269-1277) on the GPU, against the
  // edge list the graph was built from; nedge = traversed edges.
  bool validate(int64_t root, int64_t* nedge, int64_t errs[G500_VERR_NUM] = nullptr) {
    int64_t e[G500_VERR_NUM];
    int ok = 0;
    check(g500_bfs_validate(engine_, root, nullptr, nedge, e, &ok), "validate");
    if (errs)
      for (int i = 0; i < G500_VERR_NUM; ++i) errs[i] = e[i];
    return ok != 0;
  }

  // Roots as find_roots draws them (mpi/benchmark_helper.hpp:527-562).
  int find_roots(int64_t* roots, int nroots, int r1, int r2) {
    int found = 0;
    check(g500_find_roots(graph_.h, nroots, r1, r2, roots, &found), "find_roots");
    return found;
  }

  bool stats(g500_run_stats* st, std::vector<g500_level_trace>* levels = nullptr) {
    if (!engine_) return false;
    g500_level_trace tr[256];
    check(g500_bfs_get_stats(engine_, st, tr, 256), "stats");
    if (levels) levels->assign(tr, tr + (st->nlevels < 256 ? st->nlevels : 256));
    return true;
  }

  // BfsBase::printInformation (mpi/bfs.hpp:3061): settings of this engine.
  static void printInformation(int validation_level, bool pre_exec,
                               bool real_benchmark) {
    std::printf("===== Settings and Parameters. ====\n");
    std::printf("engine = MI355X direction-optimizing BFS (HIP, gfx950, 1 GPU).\n");
    std::printf("VERTEX_REORDERING = 2 (degree descending).\n");
    std::printf("ISOLATE_FIRST_EDGE = 1 (dense first-neighbour array).\n");
    std::printf("NBPE = 64.\n");
    std::printf("validation_level = %d.\n", validation_level);
    std::printf("pre_exec = %d, real_benchmark = %d.\n", (int)pre_exec, (int)real_benchmark);
  }

  GraphView graph_;

 private:
  void reset_graph() {
    end_bfs();
    if (graph_.h) g500_graph_free(graph_.h);
    graph_ = GraphView{};
  }
  void finish_construct(bool corebfs, bool symmetric) {
    const int flags = (corebfs ? G500_CONSTRUCT_COREBFS : 0) |
                      (symmetric ? G500_CONSTRUCT_SYMMETRIC_INPUT : 0);
    check(g500_graph_construct(graph_.h, flags), "construct");
    g500_graph_info inf;
    check(g500_graph_get_info(graph_.h, &inf), "graph info");
    graph_.log_orig_global_verts_ = inf.log_n;
    graph_.num_orig_local_verts_ = inf.n;
    graph_.num_global_edges_ = inf.m;
    graph_.num_global_verts_ = inf.n_c;
    graph_.max_used_vertex_ = inf.max_used_vertex;
    graph_.reorder_bits_ = inf.log_n;  // new IDs are log_n-bit wide
  }

  int device_;
  g500_bfs* engine_ = nullptr;
};

}  // namespace g500host
