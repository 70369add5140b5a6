// geometry.cpp -- host-side dual-mesh construction, the replacement of
// geometric() (src/geometric.f:2-63) behind fvm_geometric().
//
// The edge numbering and orientation are those of the reference because
// they fix the order in which the residual is summed:
//   * neighbours of a point are listed in order of first appearance while
//     scanning its triangles in ascending order, nodes 1..3 of each
//     (src/pt_surr_pt.f:20-33); the edges are the pairs (p, q > p) taken
//     point by point in that order (src/create_edge.f:9-18);
//   * an edge keeps (n1, n2) when some triangle has n1 -> n2 counter-
//     clockwise (that triangle is its left neighbour), otherwise it is
//     flipped (src/el_surr_edge.f:9-33).
// The arithmetic of the metrics follows the reference expression by
// expression (tri_coord.f, areas.f, flength.f, dtlength.f) so that the
// FP64 residual built on them can be compared bit for bit.
#include "../../include/fvm2d_gpu.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kSolid = 3;
constexpr int kFarfield = 5;

struct Pt { double x, y; };

inline double sq(double v) { return v * v; }

// parallel loop over [0, n) on the host cores
template <class F>
void parallel_for(long n, F f)
{
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) { for (long i = 0; i < n; ++i) f(i); return; }
    std::vector<std::thread> th;
    long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        long lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        th.emplace_back([=] { for (long i = lo; i < hi; ++i) f(i); });
    }
    for (auto &t : th) t.join();
}

}  // namespace

extern "C" int fvm_geometric(int np, int nt, const double *coord, int *elem, const int *ptype,
                             int cell_type, int iflow, int sa_compat, int max_edges,
                             int *edge, int *edneigh, int *iedge, double *ds, double *dsb,
                             double *elarea, double *cvarea, double *mcarea, double *drmin,
                             double *wd, double *wd1, int *beindx, int *spts, int *dims)
{
    std::memset(dims, 0, sizeof(int) * 12);
    if (np <= 0 || nt <= 0 || (cell_type != 1 && cell_type != 2)) return FVM_E_INPUT;
    const Pt *P = reinterpret_cast<const Pt *>(coord);      // P[p-1]
    auto V = [&](int t, int k) -> int & { return elem[3 * (t - 1) + k]; };

    // point classes (src/read_grid.f:62-94)
    int nsp = 0, nfp = 0;
    for (int p = 1; p <= np; ++p) {
        if (ptype[p - 1] == kSolid) spts[nsp++] = p;
        if (ptype[p - 1] == kFarfield) ++nfp;
    }

    // counter-clockwise triangles (tri_orient.f:15-41)
    for (int t = 1; t <= nt; ++t) {
        Pt a = P[V(t, 0) - 1], b = P[V(t, 1) - 1], c = P[V(t, 2) - 1];
        double cross = (b.x - a.x) * (c.y - b.y) - (c.x - b.x) * (b.y - a.y);
        if (cross == 0.0) { dims[11] = t; return FVM_E_DEGENERATE; }
        if (cross < 0.0) std::swap(V(t, 1), V(t, 2));
    }

    // triangles of each point, ascending (CSR)
    std::vector<int> t_ptr(np + 1, 0), t_idx(3 * (size_t)nt);
    for (int t = 1; t <= nt; ++t)
        for (int k = 0; k < 3; ++k) ++t_ptr[V(t, k)];
    for (int p = 1; p <= np; ++p) t_ptr[p] += t_ptr[p - 1];
    {
        std::vector<int> fill(t_ptr.begin(), t_ptr.end() - 1);
        for (int t = 1; t <= nt; ++t)
            for (int k = 0; k < 3; ++k) t_idx[fill[V(t, k) - 1]++] = t;
    }

    // neighbours in first-appearance order, then the edge list
    std::vector<int> n_ptr(np + 1, 0), n_idx;
    n_idx.reserve(6 * (size_t)np + 16);
    {
        std::vector<int> seen(np + 1, 0);
        for (int p = 1; p <= np; ++p) {
            for (int s = t_ptr[p - 1]; s < t_ptr[p]; ++s) {
                int t = t_idx[s];
                for (int k = 0; k < 3; ++k) {
                    int q = V(t, k);
                    if (q != p && seen[q] != p) { seen[q] = p; n_idx.push_back(q); }
                }
            }
            n_ptr[p] = (int)n_idx.size();
        }
    }
    // e_first[p]..: edges whose lower end is p, with their upper end
    std::vector<int> e_ptr(np + 1, 0), e_hi;
    e_hi.reserve(n_idx.size() / 2 + 1);
    for (int p = 1; p <= np; ++p) {
        for (int s = n_ptr[p - 1]; s < n_ptr[p]; ++s)
            if (n_idx[s] > p) e_hi.push_back(n_idx[s]);
        e_ptr[p] = (int)e_hi.size();
    }
    const int netot = (int)e_hi.size();
    if (netot > max_edges) return FVM_E_SIZE;
    for (int p = 1; p <= np; ++p)
        for (int e = e_ptr[p - 1]; e < e_ptr[p]; ++e) {
            edge[2 * e] = p;
            edge[2 * e + 1] = e_hi[e];
            edneigh[2 * e] = edneigh[2 * e + 1] = 0;
        }
    auto edge_of = [&](int a, int b) -> int {     // 0-based edge id of {a,b}
        int lo = std::min(a, b), hi = std::max(a, b);
        for (int e = e_ptr[lo - 1]; e < e_ptr[lo]; ++e)
            if (e_hi[e] == hi) return e;
        return -1;
    };
    // left/right triangles from the directed sides of every triangle
    for (int t = 1; t <= nt; ++t)
        for (int k = 0; k < 3; ++k) {
            int a = V(t, k), b = V(t, (k + 1) % 3);
            int e = edge_of(a, b);
            if (edge[2 * e] == a) edneigh[2 * e] = t;      // a -> b counter-clockwise in t
            else edneigh[2 * e + 1] = t;
        }
    for (int e = 0; e < netot; ++e)
        if (edneigh[2 * e] == 0) {
            std::swap(edneigh[2 * e], edneigh[2 * e + 1]);
            std::swap(edge[2 * e], edge[2 * e + 1]);
        }

    // triangle points of the dual cells: centroid or circumcentre
    std::vector<Pt> tc(nt);
    parallel_for(nt, [&](long i) {
        int t = (int)i + 1;
        Pt a = P[V(t, 0) - 1], b = P[V(t, 1) - 1], c = P[V(t, 2) - 1];
        if (cell_type == 1) {
            tc[i] = {(a.x + b.x + c.x) / 3.0, (a.y + b.y + c.y) / 3.0};
            return;
        }
        double l1 = sq(b.x - c.x) + sq(b.y - c.y);
        double l2 = sq(c.x - a.x) + sq(c.y - a.y);
        double l3 = sq(a.x - b.x) + sq(a.y - b.y);
        double det = (b.x - a.x) * (c.y - a.y) - (c.x - a.x) * (b.y - a.y);
        double b1 = 0.5 * ((b.x - a.x) * (b.x + a.x) + (b.y - a.y) * (b.y + a.y));
        double b2 = 0.5 * ((c.x - a.x) * (c.x + a.x) + (c.y - a.y) * (c.y + a.y));
        if (std::fmax(0.0, l2 + l3 - l1) == 0.0)      tc[i] = {0.5 * (b.x + c.x), 0.5 * (b.y + c.y)};
        else if (std::fmax(0.0, l3 + l1 - l2) == 0.0) tc[i] = {0.5 * (c.x + a.x), 0.5 * (c.y + a.y)};
        else if (std::fmax(0.0, l1 + l2 - l3) == 0.0) tc[i] = {0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
        else tc[i] = {((c.y - a.y) * b1 - (b.y - a.y) * b2) / det,
                      (-(c.x - a.x) * b1 + (b.x - a.x) * b2) / det};
    });

    // areas: triangles, median cells (summed per point in triangle order), CVs
    parallel_for(nt, [&](long i) {
        int t = (int)i + 1;
        Pt a = P[V(t, 0) - 1], b = P[V(t, 1) - 1], c = P[V(t, 2) - 1];
        elarea[i] = 0.5 * ((b.x - a.x) * (c.y - a.y) - (c.x - a.x) * (b.y - a.y));
    });
    // the reference stops at the first triangle (ascending) with a negative
    // sub-area (src/areas.f:109-151): the lowest one over all threads
    std::atomic<int> bad_tri{INT_MAX};
    parallel_for(np, [&](long i) {
        int p = (int)i + 1;
        double m = 0.0, cv = 0.0;
        for (int s = t_ptr[i]; s < t_ptr[i + 1]; ++s) {
            int t = t_idx[s];
            m = m + elarea[t - 1] / 3.0;
            if (cell_type == 2) {
                int j = (V(t, 0) == p) ? 0 : (V(t, 1) == p) ? 1 : 2;
                Pt a = P[p - 1], prv = P[V(t, (j + 2) % 3) - 1], nxt = P[V(t, (j + 1) % 3) - 1];
                double x[5] = {a.x, 0.5 * (a.x + nxt.x), tc[t - 1].x, 0.5 * (a.x + prv.x), a.x};
                double y[5] = {a.y, 0.5 * (a.y + nxt.y), tc[t - 1].y, 0.5 * (a.y + prv.y), a.y};
                double area = 0.0;
                for (int k = 0; k < 4; ++k) area = area + x[k] * y[k + 1] - x[k + 1] * y[k];
                if (area < 0.0) {
                    int cur = bad_tri.load(std::memory_order_relaxed);
                    while (t < cur && !bad_tri.compare_exchange_weak(cur, t, std::memory_order_relaxed)) {}
                }
                cv = cv + 0.5 * area;
            }
        }
        mcarea[i] = m;
        cvarea[i] = (cell_type == 1) ? m : cv;
    });
    if (bad_tri.load() != INT_MAX) { dims[11] = bad_tri.load(); return FVM_E_AREA; }

    // wall distance (wall_dist.f:17-94), brute force over the solid edges
    if (iflow == 3) {
        std::vector<int> sedge;
        for (int e = 0; e < netot; ++e)
            if (ptype[edge[2 * e] - 1] == kSolid && ptype[edge[2 * e + 1] - 1] == kSolid)
                sedge.push_back(e);
        parallel_for(np, [&](long i) {
            Pt x = P[i];
            double w = 1.0e10;
            for (int e : sedge) {
                Pt a = P[edge[2 * e] - 1], b = P[edge[2 * e + 1] - 1];
                double dx1 = a.x - x.x, dy1 = a.y - x.y;
                w = std::fmin(w, std::sqrt(sq(dx1) + sq(dy1)));
                double dx2 = b.x - x.x, dy2 = b.y - x.y;
                w = std::fmin(w, std::sqrt(sq(dx2) + sq(dy2)));
                double dx3 = b.x - a.x;
                double dy3 = sa_compat ? b.x - a.x : b.y - a.y;   // wall_dist.f:46 uses x
                double xi = -(dx1 * dx3 + dy1 * dy3) / (sq(dx3) + sq(dy3));
                if (xi >= 0.0 && xi <= 1.0) {
                    double xp = a.x + xi * dx3, yp = a.y + xi * dy3;
                    w = std::fmin(w, std::sqrt(sq(x.x - xp) + sq(x.y - yp)));
                }
            }
            wd[i] = (ptype[i] == kSolid) ? 0.0 : w;
        });
        for (int k = 0; k < nsp; ++k) {
            int p = spts[k];
            double w1 = 1.0e10;
            for (int s = n_ptr[p - 1]; s < n_ptr[p]; ++s)
                if (ptype[n_idx[s] - 1] != kSolid) w1 = std::fmin(w1, wd[n_idx[s] - 1]);
            if (wd1) wd1[k] = w1;
        }
        for (int p = 0; p < np; ++p)
            if (ptype[p] == kSolid) wd[p] = 1.0;
    }

    // dual-face normals (flength.f:13-41) and boundary normals (:44-61)
    bool orphan = false;
    int orphan_e = 0;
    for (int e = 0; e < netot; ++e) {
        int n1 = edge[2 * e], n2 = edge[2 * e + 1], l = edneigh[2 * e], r = edneigh[2 * e + 1];
        if (l == 0) { orphan = true; orphan_e = e + 1; break; }
        Pt a = tc[l - 1];
        Pt b = r ? tc[r - 1] : Pt{0.5 * (P[n1 - 1].x + P[n2 - 1].x), 0.5 * (P[n1 - 1].y + P[n2 - 1].y)};
        ds[2 * e] = -(b.y - a.y);
        ds[2 * e + 1] = (b.x - a.x);
    }
    if (orphan) { dims[11] = orphan_e; return FVM_E_ORPHAN_EDGE; }
    std::fill(dsb, dsb + 2 * (size_t)np, 0.0);
    for (int e = 0; e < netot; ++e) {
        if (edneigh[2 * e] != 0 && edneigh[2 * e + 1] != 0) continue;
        int n1 = edge[2 * e], n2 = edge[2 * e + 1];
        double nx = P[n2 - 1].y - P[n1 - 1].y, ny = -(P[n2 - 1].x - P[n1 - 1].x);
        dsb[2 * (n1 - 1)] += 0.5 * nx; dsb[2 * (n1 - 1) + 1] += 0.5 * ny;
        dsb[2 * (n2 - 1)] += 0.5 * nx; dsb[2 * (n2 - 1) + 1] += 0.5 * ny;
    }

    // live edges: Barth cells drop faces of zero length (flength.f:84-97)
    int ne = 0;
    for (int e = 0; e < netot; ++e)
        if (cell_type == 1 || std::sqrt(sq(ds[2 * e]) + sq(ds[2 * e + 1])) != 0.0) iedge[ne++] = e + 1;

    // time-step length: smallest altitude at each point (dtlength.f:17-44)
    std::fill(drmin, drmin + np, 1.0e15);
    for (int t = 1; t <= nt; ++t) {
        double A = elarea[t - 1];
        for (int k = 0; k < 3; ++k) {
            Pt b = P[V(t, (k + 1) % 3) - 1], c = P[V(t, (k + 2) % 3) - 1];
            double h = 2.0 * A / std::sqrt(sq(b.x - c.x) + sq(b.y - c.y));
            int p = V(t, k) - 1;
            drmin[p] = std::fmin(h, drmin[p]);
        }
    }

    // boundary edges: solid first, then farfield (make_bdedges.f:11-44)
    int nbe = 0;
    for (int pass = 0; pass < 2; ++pass) {
        int want = pass == 0 ? kSolid : kFarfield;
        for (int e = 0; e < netot; ++e) {
            if (edneigh[2 * e] != 0 && edneigh[2 * e + 1] != 0) continue;
            if (ptype[edge[2 * e] - 1] == want && ptype[edge[2 * e + 1] - 1] == want)
                beindx[nbe++] = e + 1;
        }
        if (pass == 0) { dims[7] = 1; dims[8] = nbe; dims[9] = nbe + 1; }
    }
    dims[10] = nbe;

    // area checks (geom_stat.f:18-21,34-37)
    for (int t = 0; t < nt; ++t)
        if (!(elarea[t] > 0.0)) { dims[11] = t + 1; return FVM_E_AREA; }
    for (int p = 0; p < np; ++p)
        if (!(cvarea[p] > 0.0)) { dims[11] = p + 1; return FVM_E_AREA; }

    dims[0] = np; dims[1] = nt; dims[2] = ne; dims[3] = netot;
    dims[4] = nsp; dims[5] = nfp; dims[6] = nbe;
    return FVM_OK;
}
