/*
 * g500_oracle.h -- CPU restatement of the Graph500-BFS single-rank path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under oracle/ is part of the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 *
 * Every function names the reference file:line whose behaviour it restates
 * (paths relative to the RIKEN-RCCS/Graph500-BFS tree).  Vertex IDs are the
 * reference's "original" IDs (generator output after scramble) unless a name
 * says "new" (position in the degree-descending relabel).
 */
#ifndef G500_ORACLE_H
#define G500_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- splittable MRG (generator/splittable_mrg.c) --------------------- */
typedef struct orc_mrg_state {
  uint32_t z1, z2, z3, z4, z5;
} orc_mrg_state;

/* generator/graph_generator.hpp:492-498 */
void orc_make_mrg_seed(uint64_t userseed1, uint64_t userseed2,
                       orc_mrg_state* st);
/* generator/splittable_mrg.c:224-241 (state <- A^e * state, e 192-bit) */
void orc_mrg_skip(orc_mrg_state* st, uint64_t e_high, uint64_t e_mid,
                  uint64_t e_low);
/* generator/splittable_mrg.c:331-334 */
uint32_t orc_mrg_get_uint(orc_mrg_state* st);
/* generator/splittable_mrg.c:337-343 */
double orc_mrg_get_double(orc_mrg_state* st);
/* generator/graph_generator.hpp:503-520 */
void orc_make_random_numbers(int64_t nvalues, uint64_t userseed1,
                             uint64_t userseed2, int64_t position,
                             double* result);

/* ---- Kronecker generator (generator/graph_generator.hpp) ------------- */
/* graph_generator.hpp:528-538: the two scramble words for a seed pair */
void orc_scramble_values(uint64_t userseed1, uint64_t userseed2,
                         uint64_t out[2]);
/* graph_generator.hpp:625-637 */
int64_t orc_scramble(int64_t v, int scale, uint64_t val0, uint64_t val1);
/* graph_generator.hpp:781-892 (RmatGraphGenerator<.,5700,1900>):
 * edges [start, end) of the R-MAT stream, original IDs. */
void orc_generate_edges(int scale, uint64_t userseed1, uint64_t userseed2,
                        int64_t start, int64_t end, int64_t* v0, int64_t* v1);

/* ---- graph construction (benchmark_helper.hpp, graph_constructor.hpp) */
typedef struct orc_graph orc_graph;

/* Builds the single-rank CSR the way graph500 constructs it:
 *  - self loops dropped and both directions emitted
 *    (benchmark_helper.hpp:389-497),
 *  - log2 of the vertex space = msb(OR of endpoint IDs)+1
 *    (benchmark_helper.hpp:488-496),
 *  - degree = symmetric non-loop edge count incl. duplicates and
 *    vertices relabelled by descending degree
 *    (graph_constructor.hpp:572-761; ties broken by ascending original ID,
 *     the reference's std::sort leaves ties unspecified),
 *  - rows sorted by relabelled target and duplicates merged
 *    (graph_constructor.hpp:1968-2101).
 * Returns NULL on allocation failure. */
orc_graph* orc_graph_build(const int64_t* v0, const int64_t* v1,
                           int64_t nedges);
/* Wraps CSR arrays produced elsewhere (e.g. copied back from the GPU) so the
 * BFS / schedule routines can run on them.  Arrays are not copied. */
orc_graph* orc_graph_wrap(int log_n, int64_t n_c, int64_t m,
                          const int64_t* row_off, const uint32_t* col,
                          const int64_t* deg_new, const uint32_t* new2orig,
                          const uint32_t* orig2new);
void orc_graph_free(orc_graph* g);

typedef struct orc_graph_info {
  int log_n;             /* estimated SCALE (log_orig_global_verts_) */
  int64_t n;             /* 2^log_n: pred array length at one rank */
  int64_t n_c;           /* vertices with >= 1 edge (num_global_verts_) */
  int64_t m;             /* deduplicated directed edges (num_global_edges_) */
  int64_t m_sym;         /* symmetric edges before dedupe */
  int64_t max_used_vertex; /* benchmark_helper.hpp:564-582 */
} orc_graph_info;
void orc_graph_get_info(const orc_graph* g, orc_graph_info* info);
/* Pointers into the graph (valid until orc_graph_free). */
const int64_t* orc_graph_row_off(const orc_graph* g);   /* n+1, by new ID */
const uint32_t* orc_graph_col(const orc_graph* g);      /* m, new IDs */
const int64_t* orc_graph_deg_new(const orc_graph* g);   /* n, by new ID */
const uint32_t* orc_graph_new2orig(const orc_graph* g); /* n */
const uint32_t* orc_graph_orig2new(const orc_graph* g); /* n */

/* bfs.hpp:167-170 + graph_constructor.hpp:215-245 (no CoreBFS) */
int orc_has_edge(const orc_graph* g, int64_t v);

/* ---- BFS --------------------------------------------------------------- */
/* Textbook level-synchronous BFS.  level_by_orig[v] = hop distance from
 * root, -1 if unreached.  pred_by_orig (may be NULL) receives a valid packed
 * pred entry (parent | depth<<48, -1 if unreached) as the reference writes
 * them (bfs.hpp:1457,1552,3026).  Returns the number of levels (max depth+1).
 */
int orc_bfs(const orc_graph* g, int64_t root, int32_t* level_by_orig,
            int64_t* pred_by_orig);

/* Direction-optimizing BFS, serial, following the reference's level loop
 * (bfs.hpp:3238-3566) with its alpha/beta switch (bfs.hpp:3413-3457).
 * Writes packed pred, fills at_data[9] (parameters.h:260-269) and
 * dirs[level-1] = 1 top-down / 0 bottom-up for each executed level.
 * Returns the number of executed levels (including the final empty one). */
int orc_do_bfs(const orc_graph* g, int64_t root, int edgefactor, double alpha,
               double beta, int64_t* pred_by_orig, int64_t at_data[9],
               int8_t* dirs, int max_dirs);

/* orc_do_bfs with OpenMP over nthreads threads (the reference's CPU path
 * parallelises its sweeps the same way, bfs.hpp:1355-1395, 2828-2978).  Same
 * at_data, dirs and depths; top-down parents may differ (first claimer). */
int orc_do_bfs_omp(const orc_graph* g, int64_t root, int edgefactor, double alpha,
                   double beta, int64_t* pred_by_orig, int64_t at_data[9],
                   int8_t* dirs, int max_dirs, int nthreads);

/* Replays only the alpha/beta decision logic of bfs.hpp:3298-3457 given the
 * per-level frontier sizes and frontier degree sums (index 1..nlev-1 are
 * levels 1..; index 0 unused).  Same outputs as orc_do_bfs. */
int orc_do_schedule(int64_t m, int64_t n_c, int edgefactor, double alpha,
                    double beta, int root_level, const int64_t* nq_size,
                    const int64_t* nq_edges, int nlev, int64_t at_data[9],
                    int8_t* dirs, int max_dirs);

/* ---- validation (validate.hpp:608-1255) ------------------------------- */
enum {
  ORC_ERR_RANGE = 0,      /* parent out of [-1, nglobalverts) */
  ORC_ERR_ROOT,           /* root not its own parent / depth != 0 */
  ORC_ERR_SELF_PARENT,    /* a non-root vertex is its own parent */
  ORC_ERR_DEPTH_MAP,      /* depth(v) != depth(parent)+1, unreached marks */
  ORC_ERR_EDGE_TREE,      /* edge joins a tree and a non-tree vertex */
  ORC_ERR_EDGE_GAP,       /* |depth(u)-depth(v)| > 1 */
  ORC_ERR_PRED_DEPTH,     /* check_pred_depth (validate.hpp:1018-1035) */
  ORC_ERR_NO_PARENT_EDGE, /* no input edge to the claimed parent */
  ORC_ERR_NUM
};
/* Returns 1 if valid.  te_out receives the traversed-edge count (input
 * edges with a reached source, duplicates and self loops included).
 * errs[ORC_ERR_NUM] receives per-check error counts.  With corebfs != 0 the
 * depths of pred are rebuilt from the parents first (validate.hpp:1040-1141)
 * and written back into pred, as the reference does. */
int orc_validate(const int64_t* v0, const int64_t* v1, int64_t nedges,
                 int64_t nglobalverts, int64_t nlocalverts, int corebfs,
                 int64_t root, int64_t* pred, int64_t* te_out,
                 int64_t errs[ORC_ERR_NUM]);

/* ---- CoreBFS (mpi/corebfs_adaptor.hpp, external/indexed_bfs/bfs/corebfs) */
/* Iterative leaf peeling of prune_trees (decomposition.hpp:350-572) on the
 * deduplicated symmetric graph.  Per iteration: leaves = live degree 1; each
 * leaf takes its one neighbour that is live at the iteration start as parent
 * (find_mark_twigs); then leaves drop to degree 0 (update_by_source) and each
 * parent loses one degree, dying at 0 (update_by_target).
 * Outputs by original ID: tree_parent (-1 none), has_parent (a parent was
 * assigned), in_core (live degree > 0 at the end).  stats = {core vertices,
 * core directed edges, vertices with a parent, iterations}. */
int orc_corebfs_prune(const orc_graph* g, int64_t* tree_parent,
                      uint8_t* has_parent, uint8_t* in_core, int64_t stats[4]);
/* corebfs_index::reset_unreachable (corebfs_adaptor.hpp:228-249 +
 * postprocess.hpp:106-154): BFS over the core from its highest-degree vertex
 * (ties: lowest original ID), then tree parents whose vertex cannot reach
 * that component through the forest become -1 (has_parent stays set).
 * Returns the number of parents removed; *hd_root = the BFS root. */
int64_t orc_corebfs_reset_unreachable(const orc_graph* g, int64_t* tree_parent,
                                      const uint8_t* has_parent,
                                      const uint8_t* in_core, int64_t* hd_root);
/* BfsBase::run_bfs with CoreBFS (bfs.hpp:3214-3232): walk up the forest from
 * root (bfs_tree_with_callback, corebfs.hpp:125-157), BFS inside the core from
 * the first core vertex at the walked distance, then write every tree
 * vertex's parent with depth 0xFFFF and the walked path with its depths
 * (write_tree_parents, corebfs_adaptor.hpp:285-301).  pred as the
 * reference leaves it before validation.  Returns the core root's level. */
int orc_corebfs_run(const orc_graph* g, const int64_t* tree_parent,
                    const uint8_t* has_parent, const uint8_t* in_core,
                    int64_t root, int64_t* pred);

/* ---- roots / metric ---------------------------------------------------- */
/* benchmark_helper.hpp:527-562 (single rank).  Returns the root count. */
int orc_find_roots(const orc_graph* g, int nroots, int r1, int r2,
                   int64_t* roots);
/* logfile.h:27-112: harmonic-mean TEPS of TE_i / t_i. */
double orc_harmonic_mean_teps(const double* times, const double* nedges,
                              int n);
#ifdef __cplusplus
}
#endif
#endif
