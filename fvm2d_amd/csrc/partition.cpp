// partition.cpp -- vertex patches for the LDS-tiled stage kernel.
//
// Recursive graph bisection: a vertex set is ordered by breadth-first
// distance from a pseudo-peripheral vertex of its own subgraph and cut in
// two at a multiple of the target patch size; the halves are split again
// until they fit.  On a triangle mesh this yields compact patches (small
// 1- and 2-ring halos), and numbering the vertices patch by patch keeps
// every patch's vertices contiguous in memory.
#include "partition.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace fvm {

namespace {

// Subsets are disjoint, so tasks on different subsets share the stamp array:
// each BFS draws its marks from one atomic counter, and a task only reads
// the stamp of a vertex outside its subset (relaxed atomics), never one of
// its own marks there.
struct Bisector {
    const std::vector<int> &ptr, &adj;
    std::vector<std::atomic<int>> stamp;
    std::atomic<int> epoch{0};
    int target;

    Bisector(const std::vector<int> &p, const std::vector<int> &a, int n, int t)
        : ptr(p), adj(a), stamp(n), target(t)
    {
        for (auto &s : stamp) s.store(0, std::memory_order_relaxed);
    }
    int get(int v) const { return stamp[v].load(std::memory_order_relaxed); }
    void put(int v, int m) { stamp[v].store(m, std::memory_order_relaxed); }

    // BFS restricted to the vertices stamped `mark`; returns the visit order
    // (covering every component) and leaves the last vertex reached first.
    std::vector<int> bfs(const std::vector<int> &set, int start, int mark, int visited)
    {
        std::vector<int> order;
        order.reserve(set.size());
        size_t next_seed = 0;
        int s = start;
        while (order.size() < set.size()) {
            if (get(s) != mark) {
                while (next_seed < set.size() && get(set[next_seed]) != mark) ++next_seed;
                s = set[next_seed];
            }
            // the visit order doubles as the queue
            size_t head = order.size();
            put(s, visited);
            order.push_back(s);
            while (head < order.size()) {
                int v = order[head++];
                for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
                    int w = adj[k];
                    if (get(w) == mark) { put(w, visited); order.push_back(w); }
                }
            }
        }
        return order;
    }

    // Leaves of `set` in order; the two halves of the top `par_depth` levels
    // are split on threads of their own (the leaves, and so the partition,
    // are those of the serial recursion).
    void split(std::vector<int> set, std::vector<std::vector<int>> &leaves, int par_depth)
    {
        if ((int)set.size() <= target) { leaves.push_back(std::move(set)); return; }
        int m1 = ++epoch, v1 = ++epoch;
        for (int v : set) put(v, m1);
        std::vector<int> o1 = bfs(set, set[0], m1, v1);
        int m2 = ++epoch, v2 = ++epoch;
        for (int v : set) put(v, m2);
        std::vector<int> o2 = bfs(set, o1.back(), m2, v2);   // from a far vertex
        std::vector<int>().swap(o1);
        std::vector<int>().swap(set);
        size_t blocks = (o2.size() + target - 1) / target;
        size_t cut = (size_t)target * ((blocks + 1) / 2);
        std::vector<int> a(o2.begin(), o2.begin() + cut), b(o2.begin() + cut, o2.end());
        std::vector<int>().swap(o2);
        if (par_depth > 0 && b.size() > 4096) {
            std::vector<std::vector<int>> right;
            std::thread t([&] { split(std::move(b), right, par_depth - 1); });
            split(std::move(a), leaves, par_depth - 1);
            t.join();
            for (auto &l : right) leaves.push_back(std::move(l));
        } else {
            split(std::move(a), leaves, 0);
            split(std::move(b), leaves, 0);
        }
    }
};

}  // namespace

Partition partition_vertices(int np, const std::vector<int> &adj_ptr, const std::vector<int> &adj,
                             int target)
{
    Partition p;
    std::vector<int> all(np);
    for (int v = 0; v < np; ++v) all[v] = v;
    std::vector<std::vector<int>> leaves;
    Bisector bs(adj_ptr, adj, np, std::max(1, target));
    // up to 16 threads (depth 4) on large meshes, one on small ones
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int depth = 0;
    while (depth < 4 && (2u << depth) <= hw && np > (65536 << depth)) ++depth;
    bs.split(std::move(all), leaves, depth);
    p.perm.reserve(np);
    p.patch_ptr.push_back(0);
    for (auto &l : leaves) {
        std::sort(l.begin(), l.end());            // keep the original order inside a patch
        p.perm.insert(p.perm.end(), l.begin(), l.end());
        p.patch_ptr.push_back((int)p.perm.size());
    }
    return p;
}

}  // namespace fvm
