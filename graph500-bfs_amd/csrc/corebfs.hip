// corebfs.hip -- CoreBFS forest pruning (placeholder until the device
// implementation lands).
#include "internal.hpp"

namespace g500 {

void corebfs_prune(g500_graph* g) {
  (void)g;
  fail(G500_ERR_ARG, "CoreBFS (-C) construction is not available in this build");
}

}  // namespace g500
