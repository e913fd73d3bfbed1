/*
 * g500bfs.h -- C ABI of the MI355X direction-optimizing BFS.
 *
 * Drop-in boundary for the single-rank path of RIKEN-RCCS/Graph500-BFS.
 * The reference has no runtime plugin interface: its seam is the compile-time
 * typedef `BfsOnCPU` (mpi/bfs_cpu.hpp:13) and the driver calls the BfsBase
 * method set (mpi/bfs.hpp:116-172).  Each entry point below names the
 * reference interface it replaces.  The C++ class `BfsOnMI355X`
 * (graph500-bfs_amd/host/bfs_mi355x.hpp) wraps these calls with exactly that
 * method set.
 *
 * Conventions
 *  - Every call returns G500_OK (0) or a negative G500_ERR_* code; the message
 *    of the last failure on the calling thread is g500_last_error().
 *  - Host buffers belong to the caller; device buffers belong to the handle.
 *  - Vertex IDs at the boundary are the generator's ("original") IDs, as in
 *    the reference.  A pred entry packs parent (bits 0-47, signed) and BFS
 *    depth (bits 48-63, 0xFFFF = unreached); -1 = unreached
 *    (mpi/validate.hpp:982-992, mpi/bfs.hpp:1457).
 *  - One BFS at a time per g500_bfs handle (as BfsBase); handles on the same
 *    graph may run concurrently on their own streams.
 */
#ifndef G500BFS_H
#define G500BFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  G500_OK = 0,
  G500_ERR_ARG = -1,     /* invalid argument */
  G500_ERR_HIP = -2,     /* HIP runtime failure */
  G500_ERR_NOMEM = -3,   /* device or host allocation failed */
  G500_ERR_STATE = -4,   /* call out of order (e.g. run before construct) */
  G500_ERR_NODEV = -5    /* no usable GPU */
};

/* auto_tuning_data slots, mpi/parameters.h:260-269 */
enum {
  G500_AT_GLOBAL_NQ_EDGES = 0,
  G500_AT_NUM_GLOBAL_EDGES = 1,
  G500_AT_T2B_LEVEL = 2,
  G500_AT_GLOBAL_NQ_SIZE = 3,
  G500_AT_NUM_GLOBAL_VERTS = 4,
  G500_AT_B2T_LEVEL = 5,
  G500_AT_PRE_GLOBAL_NQ_EDGES = 6,
  G500_AT_PRE_GLOBAL_NQ_SIZE = 7,
  G500_AT_LEVEL = 8,
  G500_AT_NUM = 9
};

/* validation error classes, in the order mpi/validate.hpp:608-976 checks them */
enum {
  G500_VERR_RANGE = 0,
  G500_VERR_ROOT = 1,
  G500_VERR_SELF_PARENT = 2,
  G500_VERR_DEPTH_MAP = 3,
  G500_VERR_EDGE_TREE = 4,
  G500_VERR_EDGE_GAP = 5,
  G500_VERR_PRED_DEPTH = 6,
  G500_VERR_NO_PARENT_EDGE = 7,
  G500_VERR_NUM = 8
};

typedef struct g500_graph g500_graph;
typedef struct g500_bfs g500_bfs;

typedef struct g500_graph_info {
  int log_n;               /* estimated SCALE, Graph2DCSR::log_orig_global_verts_ */
  int corebfs;             /* 1 if constructed with CoreBFS pruning (-C) */
  int64_t n;               /* 2^log_n = pred length, num_orig_local_verts_ */
  int64_t n_c;             /* vertices with >= 1 CSR edge, num_global_verts_ */
  int64_t m;               /* deduplicated directed CSR edges, num_global_edges_ */
  int64_t m_sym;           /* symmetric edges before dedupe */
  int64_t nedges_input;    /* generated / uploaded edges (validation input) */
  int64_t max_used_vertex; /* find_max_used_vertex, benchmark_helper.hpp:564-582 */
  int64_t max_degree;      /* largest symmetric degree */
  int64_t core_vertices;   /* CoreBFS: vertices kept in the 2-core, else 0 */
  int64_t tree_vertices;   /* CoreBFS: pruned vertices holding a parent */
  int64_t device_bytes;    /* device memory held by the graph */
  int64_t removed_parents; /* CoreBFS: parents reset as unreachable (after the
                              first g500_bfs_create), else 0 */
  int64_t hd_root;         /* CoreBFS: highest-degree core vertex, else -1 */
} g500_graph_info;

/* ---- graph (L3/L4: generation + construction) -------------------------- */

/* Kronecker (R-MAT A/B/C/D = .57/.19/.19/.05) edge list generated on the GPU.
 * Replaces generate_graph_spec2010 (mpi/benchmark_helper.hpp:234-241 ->
 * generator/graph_generator.hpp:781-892).  seed1/seed2 = PRM::USERSEED1/2
 * (2, 3 in the reference, mpi/parameters.h:204-205). */
int g500_graph_generate(int device, int scale, int edgefactor, uint64_t seed1,
                        uint64_t seed2, g500_graph** out);

/* Takes an existing edge list (original IDs, < 2^32) instead of generating
 * one; nedges may be 0.  Replaces filling EdgeListStorage
 * (generator/graph_generator.hpp:46-485) from another source. */
int g500_graph_from_edges(int device, const int64_t* v0, const int64_t* v1,
                          int64_t nedges, g500_graph** out);

/* An empty graph on `device` whose input edge list will hold up to
 * `capacity` edges, filled chunk by chunk with g500_graph_append_edges and
 * then built with g500_graph_construct.  This is how BfsBase::construct
 * (mpi/bfs.hpp:116-139) hands the driver's EdgeListStorage to the device: its
 * chunks (beginRead/read/endRead, generator/graph_generator.hpp:46-485) go
 * straight to device memory sized from num_local_edges(), with no host copy
 * of the whole list. */
int g500_graph_create(int device, int64_t capacity, g500_graph** out);

/* Appends count edges (original IDs, < 2^32) after those already held;
 * fails with G500_ERR_ARG past the capacity and G500_ERR_STATE after
 * construction. */
int g500_graph_append_edges(g500_graph* g, const int64_t* v0, const int64_t* v1,
                            int64_t count);

/* flags of g500_graph_construct */
enum {
  G500_CONSTRUCT_COREBFS = 1,        /* prune trees first (-C) */
  G500_CONSTRUCT_SYMMETRIC_INPUT = 2 /* edges already hold both directions */
};

/* Symmetrize, relabel by degree and build the CSR on the GPU; with
 * G500_CONSTRUCT_COREBFS also prune trees (CoreBFS) first.  Replaces
 * make_symmetry_edge_list (mpi/benchmark_helper.hpp:389-497) +
 * BfsBase::construct (mpi/bfs.hpp:116-139).  With
 * G500_CONSTRUCT_SYMMETRIC_INPUT the edge list is taken as the output of
 * make_symmetry_edge_list (what BfsBase::construct receives): each directed
 * edge counts once toward its source's degree. */
int g500_graph_construct(g500_graph* g, int flags);

int g500_graph_get_info(const g500_graph* g, g500_graph_info* info);

/* BfsBase::has_edge (mpi/bfs.hpp:167-170). */
int g500_graph_has_edge(g500_graph* g, int64_t v, int* out);

/* Copies input edges [start, start+count) to host int64 arrays. */
int g500_graph_copy_edges(const g500_graph* g, int64_t start, int64_t count,
                          int64_t* v0, int64_t* v1);

/* Copies the constructed CSR to the host (any pointer may be NULL).
 * row_off: n+1 int64 by new ID; col: m uint32 new IDs; deg_new: n int64
 * symmetric degree by new ID; new2orig / orig2new: n uint32. */
int g500_graph_copy_csr(const g500_graph* g, int64_t* row_off, uint32_t* col,
                        int64_t* deg_new, uint32_t* new2orig,
                        uint32_t* orig2new);

/* CoreBFS forest by original ID: tree_parent (n, -1 = none or reset) and
 * has_parent (n bytes, 1 = a parent was assigned while pruning; has_edge
 * counts these vertices).  Either pointer may be NULL. */
int g500_graph_copy_forest(const g500_graph* g, int64_t* tree_parent,
                           uint8_t* has_parent);

/* Roots drawn as find_roots does (mpi/benchmark_helper.hpp:527-562) with
 * seeds (seed1 + r1, seed2 + r2); returns the count in *found. */
int g500_find_roots(g500_graph* g, int nroots, int r1, int r2, int64_t* roots,
                    int* found);

/* make_random_numbers (generator/graph_generator.hpp:503-520): nvalues
 * doubles from the MRG stream of (seed1, seed2) starting at `position`
 * (the driver's -S root-seed search uses it). */
int g500_make_random_numbers(int64_t nvalues, uint64_t seed1, uint64_t seed2,
                             int64_t position, double* out);

/* Frees the graph.  Every engine created on it (g500_bfs, g500_msbfs) must be
 * destroyed first: with engines still alive the call fails with
 * G500_ERR_STATE and frees nothing. */
int g500_graph_free(g500_graph* g);

/* ---- BFS engine (L5) --------------------------------------------------- */

/* BfsBase::prepare_bfs -> allocate_memory (mpi/bfs.hpp:141-151, 190-303);
 * with CoreBFS also corebfs_index::reset_unreachable
 * (mpi/corebfs_adaptor.hpp:228-249). */
int g500_bfs_create(g500_graph* g, g500_bfs** out);

/* The same with prepare_bfs's edgefactor, alpha and beta (mpi/bfs.hpp:141-151),
 * which the CoreBFS reachability BFS of a graph's first engine runs with
 * (reset_unreachable, mpi/corebfs_adaptor.hpp:228-249); g500_bfs_create uses
 * the reference defaults 16, 10, 14.  Both read, at creation, G500_EXEC_ALPHA
 * (see g500_bfs_set_exec_alpha), G500_SPECULATE (guessed steps the host
 * queues ahead of the running one, default 1, 0 = none) and G500_GUESS
 * ("wrong": a test setting that makes every guess wrong; results stay the
 * same, runs get slower). */
int g500_bfs_create_ex(g500_graph* g, int edgefactor, double alpha, double beta,
                       g500_bfs** out);

/* BfsBase::run_bfs(root, pred, ef, alpha, beta, auto_tuning_data)
 * (mpi/bfs.hpp:3214-3232).  pred stays on the device until
 * g500_bfs_get_pred; at_data may be NULL.  Blocks until the BFS finished. */
int g500_bfs_run(g500_bfs* b, int64_t root, int edgefactor, double alpha,
                 double beta, int64_t* at_data);

/* BfsBase::get_pred (mpi/bfs.hpp:161-163): device pred -> host, n entries. */
int g500_bfs_get_pred(g500_bfs* b, int64_t* pred);

/* Depth of every vertex by original ID (pred >> 48), n entries. */
int g500_bfs_get_levels(g500_bfs* b, uint16_t* levels);

/* validate_bfs_result (mpi/validate.hpp:1269-1277) on the device against
 * the resident input edge list, using the last run's pred (or, when
 * host_pred != NULL, that array uploaded first).  nglobalverts is
 * max_used_vertex+1 as in mpi/main.cc:646.  Fills te (traversed edges),
 * errs[G500_VERR_NUM] and *ok.  With CoreBFS the rebuilt depths are written
 * back into the device pred, as the reference does in place. */
int g500_bfs_validate(g500_bfs* b, int64_t root, const int64_t* host_pred,
                      int64_t* te, int64_t* errs, int* ok);

/* Per-level trace of the last run. */
typedef struct g500_level_trace {
  int level;        /* BFS level number (root = level 0 of the core BFS) */
  int top_down;     /* direction the reference's switch chose: 1 top-down */
  int64_t nq_size;  /* vertices discovered at this level */
  int64_t nq_edges; /* sum of their degrees (top-down levels only) */
  float gpu_ms;     /* device time of the level, 0 unless timing is on */
  int exec_bottom_up; /* 1 if the level ran as a bottom-up sweep */
} g500_level_trace;

typedef struct g500_run_stats {
  int nlevels;          /* executed levels (incl. the final empty one) */
  double gpu_ms;        /* device time of the whole run (events on stream) */
  double td_ms, bu_ms;  /* device time in top-down / bottom-up levels */
  int64_t td_edges;     /* edges scanned by top-down levels */
  int64_t root_level;   /* level the core BFS started at (CoreBFS) */
  /* per-kernel device time from events on the kernel's own stream */
  double gather_ms;     /* pred gather by original ID (k_pred_gather_core / k_pred_gather) */
  double sweep_ms;      /* bottom-up bitmap sweeps (k_bu_sweep), summed over the run */
  int32_t sweeps;       /* k_bu_sweep launches in the run */
  int32_t counted;      /* 1: the run counted its algorithmic bytes (below) */
  /* Counting runs only (g500_bfs_set_counting): bytes the run's kernels need
   * to move -- each stream read or written once, the row entries a bottom-up
   * scan reads up to its first visited neighbour (counted on the device),
   * visited-bitmap probes excluded -- in total and for the sweep and gather
   * kernels (their launches in this run). */
  int64_t alg_bytes;
  int64_t sweep_alg_bytes;
  int64_t gather_alg_bytes;
  /* with level timing on (g500_bfs_set_timing): device time between the end
   * of one step's kernels (a level, or the bitmap -> queue conversion at the
   * bottom-up -> top-down switch) and the start of the next one's, summed
   * over the run's step boundaries */
  double gap_ms;
  int32_t gaps;         /* step boundaries summed in gap_ms */
  /* steps queued on a predicted direction that the device's decision did
   * not take (each returned at once; the right step was queued after the
   * level's report) */
  int32_t mispredicted;
} g500_run_stats;

/* Enables per-level HIP-event timing (off by default; costs one event pair
 * per level). */
int g500_bfs_set_timing(g500_bfs* b, int on);

/* Execution choice for top-down levels.  The direction decision and
 * auto_tuning_data always follow the reference's alpha/beta rule; a level
 * the rule runs top-down is executed as a bottom-up sweep when its frontier
 * holds more than m / exec_alpha edges (same vertices, same depths, any valid
 * parents).  0 disables (every top-down level runs top-down).  The default
 * comes from G500_EXEC_ALPHA in the environment, else 25 (on MI355X a
 * bottom-up sweep beats the top-down expansion of such levels). */
int g500_bfs_set_exec_alpha(g500_bfs* b, double exec_alpha);
/* Counting runs (on != 0): the bottom-up kernels also tally the bytes their
 * row scans need and g500_bfs_get_stats reports the run's algorithmic bytes.
 * Same results; off by default (it costs a copy after each run). */
int g500_bfs_set_counting(g500_bfs* b, int on);
int g500_bfs_get_stats(g500_bfs* b, g500_run_stats* st,
                       g500_level_trace* levels, int max_levels);

/* BfsBase::end_bfs (mpi/bfs.hpp:165): frees the engine. */
int g500_bfs_destroy(g500_bfs* b);

/* ---- Batched bit-parallel multi-source BFS -------------------------------
 * No counterpart in the reference, which runs one root per run_bfs
 * (mpi/bfs.hpp:3214).  Up to G500_MSBFS_MAX_ROOTS roots traverse the graph
 * together as 64-bit lane masks per vertex.  The result per root is its uint8
 * level array by original ID (0xFF = unreached), equal to the depths that
 * g500_bfs_run finds for that root.  Graphs built with
 * G500_CONSTRUCT_COREBFS are not supported (their CSR holds the core only). */
#define G500_MSBFS_MAX_ROOTS 256
typedef struct g500_msbfs g500_msbfs;

/* Work split of the level sweeps (new IDs are degree-descending, so each
 * degree bin is a contiguous range).  Vertices with degree >= heavy_degree are
 * cut into work items of heavy_chunk edges, one wave per item; those with
 * degree >= medium_degree get medium_group lanes each; the rest one lane
 * each.  A level runs top-down (push from the frontier) while its frontier's
 * summed degree is below m / td_alpha, else bottom-up.  While fewer than
 * stream_frac of all (root, vertex) pairs have been reached, a bottom-up level
 * is "dense" (early exit rarely pays) and runs as dense_algo says: 1 = one
 * balanced stream over every edge; 2 = every vertex pulls from its neighbours
 * below the hub cut-off (the lowest IDs, whose lane words take hot_bytes and
 * stay cached), and the frontier vertices above it push along their rows.
 * Later bottom-up levels pull per vertex with early exit; at the first of
 * them the roots with light frontiers (split_bp) push along their frontier
 * rows instead, so a vertex's pull stops once the other roots are covered.
 * blocks_per_cu sizes the grid of the persistent sweep kernels; use_graph = 1
 * replays the whole batch as one captured HIP graph. */
typedef struct g500_msbfs_tuning {
  int64_t heavy_degree;
  int64_t medium_degree;
  int32_t medium_group;  /* 2, 4, 8, 16 or 32 */
  int32_t blocks_per_cu; /* 1..32 */
  int32_t use_graph;     /* 0 or 1 */
  int32_t heavy_chunk;   /* 256..65536, multiple of 256 */
  double td_alpha;       /* > 0 */
  double stream_frac;    /* 0..1; 0 = no dense levels */
  int64_t hot_bytes;     /* dense_algo 2: lane words of the lowest IDs kept hot */
  int32_t dense_algo;    /* dense levels: 1 streaming pull, 2 hub pull + push */
  int32_t split_bp;      /* 0..10000: at the first bottom-up level after top-down
                            ones, the roots whose frontier holds fewer than
                            split_bp / 10000 * m edges (times 2/K for K > 2
                            lane words) push instead of being pulled
                            (0 = every root pulled) */
} g500_msbfs_tuning;

typedef struct g500_msbfs_stats {
  int32_t nroots;
  int32_t nlevels;        /* levels with a non-empty frontier, incl. level 0 */
  double gpu_ms;          /* device time of the last batch (events on its stream) */
  int64_t discovered;     /* (root, vertex) pairs reached, summed over roots */
  int64_t lane_discoveries[64]; /* per level (first 64 levels) */
} g500_msbfs_stats;

/* flags of g500_msbfs_create_ex / g500_msbfs_footprint */
enum {
  G500_MSBFS_NO_LEVEL_ARRAYS = 1  /* answer point lookups only: no [roots][N] arrays */
};
/* Device bytes g500_msbfs_create_ex(g, max_roots, flags) will allocate
 * (admission control of a serving loop checks it against g500_device_memory). */
int g500_msbfs_footprint(const g500_graph* g, int max_roots, int flags, int64_t* bytes);
/* Free and total device memory (hipMemGetInfo). */
int g500_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes);
/* Allocates the engine for batches of up to max_roots roots (1..256). */
int g500_msbfs_create(g500_graph* g, int max_roots, g500_msbfs** out);
/* The same with flags (G500_MSBFS_NO_LEVEL_ARRAYS: no per-root level arrays;
 * g500_msbfs_get_levels and g500_msbfs_device_levels then fail, point lookups
 * work). */
int g500_msbfs_create_ex(g500_graph* g, int max_roots, int flags, g500_msbfs** out);
int g500_msbfs_set_tuning(g500_msbfs* b, const g500_msbfs_tuning* t);
int g500_msbfs_get_tuning(g500_msbfs* b, g500_msbfs_tuning* t);
/* Enqueues a batch (roots by original ID, 1..max_roots of them, duplicates
 * allowed) on the engine's stream and returns without waiting. */
int g500_msbfs_launch(g500_msbfs* b, const int64_t* roots, int nroots);
/* *done = 1 when the launched batch has finished on the device. */
int g500_msbfs_query(g500_msbfs* b, int* done);
/* Waits for the launched batch and completes it (levels beyond the captured
 * level count run here).  Results stay on the device until read. */
int g500_msbfs_wait(g500_msbfs* b);
/* g500_msbfs_launch + g500_msbfs_wait. */
int g500_msbfs_run(g500_msbfs* b, const int64_t* roots, int nroots);
/* Level array of root `index` of the last batch: N = 2^log_n bytes. */
int g500_msbfs_get_levels(g500_msbfs* b, int index, uint8_t* host_levels);
/* Device address of the level arrays ([nroots][N] bytes, row stride N). */
int g500_msbfs_device_levels(g500_msbfs* b, void** dev_ptr, int64_t* row_stride);
/* Point lookups into the last batch: levels[i] = level of vertex pairs[2i+1]
 * (original ID) for the root at batch index pairs[2i] (0xFF = unreached or
 * vertex out of range); read from the level frontiers, so they work with or
 * without level arrays. */
int g500_msbfs_gather_levels(g500_msbfs* b, const int64_t* pairs, int64_t count,
                             uint8_t* levels);
/* The same point lookups without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * pairs and levels are staged in a file-scope device array of
 * G500_NOALLOC_MAX_PAIRS entries (about 1.1 MiB) that each GPU holds until the
 * process ends; larger counts fail with G500_ERR_ARG.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (from several host
 * threads) race. */
#define G500_NOALLOC_MAX_PAIRS 65536
int g500_msbfs_gather_levels_noalloc(g500_msbfs* b, const int64_t* pairs, int64_t count,
                                     uint8_t* levels);
/* Picks the tuning fields by timing batches of `roots` (coordinate search,
 * `rounds` passes of 1..8); leaves the best tuning set and returns it with its
 * device time per batch. */
int g500_msbfs_autotune(g500_msbfs* b, const int64_t* roots, int nroots, int rounds,
                        g500_msbfs_tuning* best, double* best_ms);
/* Traversed edges per root (input edges whose source the root reaches, the
 * validator's TE, mpi/validate.hpp:840-872); te holds nroots entries. */
int g500_msbfs_get_te(g500_msbfs* b, int64_t* te);
int g500_msbfs_get_stats(g500_msbfs* b, g500_msbfs_stats* st);
int g500_msbfs_destroy(g500_msbfs* b);

/* Message of the last failing call on this thread ("" if none). */
const char* g500_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* G500BFS_H */
