// Drives BfsOnMI355X exactly as the reference driver drives BfsBase
// (mpi/main.cc:433-646): construct(scale, corebfs, &sym_edge_list) from an
// EdgeListStorage-shaped container, prepare_bfs, run_bfs, get_pred.
// usage: dropin_main <sym_edges.bin (int64 v0,v1 pairs)> <root> <pred_out.bin>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../graph500-bfs_amd/host/bfs_mi355x.hpp"

// The read protocol of generator/graph_generator.hpp:46-485, in memory.
struct PackedEdge {
  int64_t a, b;
  int64_t v0() const { return a; }
  int64_t v1() const { return b; }
};
struct MemEdgeList {
  typedef PackedEdge edge_type;
  // small chunks: the drop-in appends each one to the device separately
  static const int CHUNK_SIZE = 1 << 16;
  std::vector<PackedEdge> e;
  size_t pos = 0;
  int64_t num_local_edges() const { return (int64_t)e.size(); }
  int beginRead(bool) {
    pos = 0;
    return (int)((e.size() + CHUNK_SIZE - 1) / CHUNK_SIZE);
  }
  int read(PackedEdge** buf) {
    *buf = e.data() + pos;
    const size_t n = std::min<size_t>(CHUNK_SIZE, e.size() - pos);
    pos += n;
    return (int)n;
  }
  void endRead() {}
};

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  MemEdgeList list;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t ab[2];
  while (std::fread(ab, sizeof(int64_t), 2, f) == 2) list.e.push_back({ab[0], ab[1]});
  std::fclose(f);
  const int64_t root = std::atoll(argv[2]);
  g500host::BfsOnMI355X bfs(0);
  bfs.construct(0, false, &list);
  std::vector<int64_t> pred((size_t)bfs.graph_.pred_size());
  bfs.prepare_bfs(2, false, false, 16, 10.0, 14.0, pred.data());
  int64_t at[9];
  for (auto& x : at) x = -1;
  bfs.run_bfs(root, pred.data(), 16, 10.0, 14.0, at);
  bfs.get_pred(pred.data());
  bfs.end_bfs();
  FILE* o = std::fopen(argv[3], "wb");
  std::fwrite(pred.data(), sizeof(int64_t), pred.size(), o);
  std::fclose(o);
  std::printf("n=%lld m=%lld nc=%lld level=%lld t2b=%lld\n",
              (long long)bfs.graph_.pred_size(), (long long)bfs.graph_.num_global_edges_,
              (long long)bfs.graph_.num_global_verts_, (long long)at[8], (long long)at[2]);
  return 0;
}
