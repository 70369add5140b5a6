// partition.h -- vertex patches for the LDS-tiled stage kernel.
#pragma once
#include <vector>

namespace fvm {

struct Partition {
    std::vector<int> perm;       // new vertex number -> original (0-based)
    std::vector<int> patch_ptr;  // patch k = new vertices [patch_ptr[k], patch_ptr[k+1])
};

// Splits the vertex graph (CSR adjacency, 0-based) into patches of at most
// `target` vertices by recursive breadth-first bisection.
Partition partition_vertices(int np, const std::vector<int> &adj_ptr, const std::vector<int> &adj,
                             int target);

}  // namespace fvm
