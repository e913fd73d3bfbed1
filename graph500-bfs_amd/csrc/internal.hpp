// internal.hpp -- handle types behind the C ABI and the entry points shared
// between the kernel files.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "primitives.hpp"

struct g500_graph {
  int device = 0;
  hipStream_t stream = nullptr;

  // input edge list (original IDs), kept resident for validation
  g500::DevBuf<uint2> edges;
  int64_t nedges = 0;
  int64_t capacity = 0;  // edges the buffer holds (appended graphs)
  uint64_t id_or = 0;    // OR of every input ID appended so far
  int id_bits = 0;  // every input ID < 2^id_bits
  bool symmetric_input = false;  // edges already hold both directions

  // construction results (all CSR arrays indexed by relabelled "new" ID)
  bool constructed = false;
  int corebfs = 0;
  int log_n = 0;       // estimated SCALE
  int64_t n = 0;       // 2^log_n
  int64_t n_c = 0;     // vertices with >= 1 CSR edge (new IDs [0, n_c))
  int64_t m = 0;       // deduplicated directed edges
  int64_t m_sym = 0;   // symmetric edges before dedupe
  int64_t max_used = 0;
  int64_t max_deg = 0;
  g500::DevBuf<int64_t> row_off;    // n_c + 1
  g500::DevBuf<uint32_t> col;       // m
  g500::DevBuf<uint2> first;        // n_c: first (highest-degree) neighbour: new ID, original ID
  g500::DevBuf<uint64_t> multi;     // n_c bits: the row has more than one entry
  g500::DevBuf<uint32_t> deg_new;   // n   (symmetric degree incl. duplicates)
  g500::DevBuf<uint32_t> new2orig;  // n
  g500::DevBuf<uint32_t> orig2new;  // n

  // host copy of orig2new (lazily made for has_edge / root selection)
  std::vector<uint32_t> h_orig2new;

  // CoreBFS (corebfs != 0): forest parents by original ID (kNoParent = none
  // or reset as unreachable) and the bitmap of vertices that were given a
  // parent while pruning (has_edge counts them, mpi/bfs.hpp:167-170)
  g500::DevBuf<uint32_t> tree_parent;     // n
  g500::DevBuf<uint64_t> has_parent;      // n/64 words
  std::vector<uint32_t> h_tree_parent;    // host copy for the per-root walk
  std::vector<uint64_t> h_has_parent;
  bool core_reset_done = false;           // reset_unreachable applied
  int64_t core_vertices = 0, tree_vertices = 0, removed_parents = 0;
  int64_t hd_root = -1;                   // highest-degree core vertex

  g500::ScanWorkspace scan_ws;
  g500::RadixWorkspace radix_ws;

  // engines (plain and batched) created on this graph and not yet destroyed:
  // they hold pointers into it, so the graph may only be freed after them
  int live_engines = 0;

  int64_t device_bytes() const;
  ~g500_graph();
};

namespace g500 {

// gen.hip
void generate_edges(g500_graph* g, int scale, int edgefactor, uint64_t seed1,
                    uint64_t seed2);

// construct.hip
// CSR of `edges` into g (force_log_n < 0: estimate the vertex space).
void build_csr(g500_graph* g, const uint2* edges, int64_t ne, bool symmetric,
               int force_log_n);
void construct_graph(g500_graph* g, int corebfs);
void ensure_host_orig2new(g500_graph* g);

// corebfs.hip
constexpr uint32_t kNoParent = 0xFFFFFFFFu;
// entries allocated past the end of col (aligned 16-byte loads of a row's
// last group may read up to three of them; their values are never used)
constexpr int64_t kColPad = 16;
void corebfs_prune(g500_graph* g);
// reachability reset of the forest after a BFS from hd_root over the core;
// `visited` is that BFS's visited bitmap over core new IDs
void corebfs_reset_unreachable(g500_graph* g, const uint64_t* visited,
                               hipStream_t s);
inline bool has_parent_host(const g500_graph* g, int64_t v) {
  return g->corebfs && (g->h_has_parent[(size_t)(v >> 6)] >> (v & 63)) & 1ull;
}

}  // namespace g500
