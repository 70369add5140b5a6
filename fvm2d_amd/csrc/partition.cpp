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
#include <deque>
#include <vector>

namespace fvm {

namespace {

struct Bisector {
    const std::vector<int> &ptr, &adj;
    std::vector<int> stamp, dist;
    int epoch = 0;
    int target;

    Bisector(const std::vector<int> &p, const std::vector<int> &a, int n, int t)
        : ptr(p), adj(a), stamp(n, 0), dist(n, 0), target(t) {}

    // BFS restricted to the vertices stamped `mark`; returns the visit order
    // (covering every component) and leaves the last vertex reached first.
    std::vector<int> bfs(const std::vector<int> &set, int start, int mark, int visited)
    {
        std::vector<int> order;
        order.reserve(set.size());
        size_t next_seed = 0;
        int s = start;
        while (order.size() < set.size()) {
            if (stamp[s] != mark) {
                while (next_seed < set.size() && stamp[set[next_seed]] != mark) ++next_seed;
                s = set[next_seed];
            }
            std::deque<int> q{s};
            stamp[s] = visited;
            while (!q.empty()) {
                int v = q.front();
                q.pop_front();
                order.push_back(v);
                for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
                    int w = adj[k];
                    if (stamp[w] == mark) { stamp[w] = visited; q.push_back(w); }
                }
            }
        }
        return order;
    }

    void split(std::vector<int> set, std::vector<std::vector<int>> &leaves)
    {
        if ((int)set.size() <= target) { leaves.push_back(std::move(set)); return; }
        int m1 = ++epoch, v1 = ++epoch;
        for (int v : set) stamp[v] = m1;
        std::vector<int> o1 = bfs(set, set[0], m1, v1);
        int m2 = ++epoch, v2 = ++epoch;
        for (int v : set) stamp[v] = m2;
        std::vector<int> o2 = bfs(set, o1.back(), m2, v2);   // from a far vertex
        size_t blocks = (set.size() + target - 1) / target;
        size_t cut = (size_t)target * ((blocks + 1) / 2);
        std::vector<int> a(o2.begin(), o2.begin() + cut), b(o2.begin() + cut, o2.end());
        split(std::move(a), leaves);
        split(std::move(b), leaves);
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
    bs.split(std::move(all), leaves);
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
