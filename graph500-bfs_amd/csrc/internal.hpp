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
  g500::DevBuf<uint32_t> first;     // n_c: first (highest-degree) neighbour
  g500::DevBuf<uint32_t> deg_new;   // n   (symmetric degree incl. duplicates)
  g500::DevBuf<uint32_t> new2orig;  // n
  g500::DevBuf<uint32_t> orig2new;  // n

  // host copy of orig2new (lazily made for has_edge / root selection)
  std::vector<uint32_t> h_orig2new;

  // CoreBFS: pruned-forest parents (original IDs, -1 = in core / none),
  // valid when corebfs != 0
  g500::DevBuf<int64_t> tree_parent;  // n
  std::vector<int64_t> h_tree_parent;
  int64_t core_vertices = 0, tree_vertices = 0;

  g500::ScanWorkspace scan_ws;
  g500::RadixWorkspace radix_ws;

  int64_t device_bytes() const;
  ~g500_graph();
};

namespace g500 {

// gen.hip
void generate_edges(g500_graph* g, int scale, int edgefactor, uint64_t seed1,
                    uint64_t seed2);

// construct.hip
void construct_graph(g500_graph* g, int corebfs);
void ensure_host_orig2new(g500_graph* g);

// corebfs.hip
void corebfs_prune(g500_graph* g);

}  // namespace g500
