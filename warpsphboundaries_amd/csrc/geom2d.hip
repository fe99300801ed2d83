// Static 2-D geometry (edge cell list, cell indicator, y-bins) and the fixed-capacity adjacency kernel.
//
// Bit-exactness contract: every floating-point operation of the slot predicate and of the cell / bin
// index arithmetic is written in the reference's order with no fused multiply-add (contract(off) for this
// whole file), so the slots (set and order) equal those of scene/fixedadj.py:67-97 and of the cell list
// built by scene/scene.py:100-119,273-279.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "cellsort.h"
#include "geom2d_device.h"
#include "wsb_internal.h"

namespace {

// ------------------------------------------------------------------------------------------------ host build
struct HostCells {
    double lo0, lo1, cell;
    int nx, ny;
    std::vector<int32_t> start, items;
};

// every edge into every cell overlapped by its support-inflated box, items of a cell in ascending edge id
// (reference scene/scene.py:100-119 buildCellList with the boxes of :273-279)
HostCells build_cells(const double* V, const int32_t* E, int ne, double smax) {
    std::vector<double> lo(2 * (size_t)ne), hi(2 * (size_t)ne);
    double g0 = INFINITY, g1 = INFINITY, G0 = -INFINITY, G1 = -INFINITY;
    for (int e = 0; e < ne; ++e) {
        const double ax = V[2 * E[2 * e]], ay = V[2 * E[2 * e] + 1];
        const double bx = V[2 * E[2 * e + 1]], by = V[2 * E[2 * e + 1] + 1];
        lo[2 * e] = std::min(ax, bx) - smax;
        lo[2 * e + 1] = std::min(ay, by) - smax;
        hi[2 * e] = std::max(ax, bx) + smax;
        hi[2 * e + 1] = std::max(ay, by) + smax;
        g0 = std::min(g0, lo[2 * e]);
        g1 = std::min(g1, lo[2 * e + 1]);
        G0 = std::max(G0, hi[2 * e]);
        G1 = std::max(G1, hi[2 * e + 1]);
    }
    HostCells c;
    c.lo0 = g0;
    c.lo1 = g1;
    c.cell = smax;
    const int64_t nx = std::max<int64_t>((int64_t)std::ceil((G0 - g0) / smax), 1);
    const int64_t ny = std::max<int64_t>((int64_t)std::ceil((G1 - g1) / smax), 1);
    c.nx = (int)nx;
    c.ny = (int)ny;
    std::vector<int64_t> cid;
    std::vector<int32_t> item;
    for (int e = 0; e < ne; ++e) {
        const int64_t i0x = std::max<int64_t>((int64_t)std::floor((lo[2 * e] - g0) / smax), 0);
        const int64_t i0y = std::max<int64_t>((int64_t)std::floor((lo[2 * e + 1] - g1) / smax), 0);
        const int64_t i1x = std::min<int64_t>((int64_t)std::floor((hi[2 * e] - g0) / smax), nx - 1);
        const int64_t i1y = std::min<int64_t>((int64_t)std::floor((hi[2 * e + 1] - g1) / smax), ny - 1);
        const int64_t cx = std::max<int64_t>(i1x - i0x + 1, 0), cy = std::max<int64_t>(i1y - i0y + 1, 0);
        for (int64_t l = 0; l < cx * cy; ++l) {
            cid.push_back((i0x + l % cx) * ny + (i0y + l / cx));
            item.push_back(e);
        }
    }
    // stable counting sort by cell id
    c.start.assign((size_t)(nx * ny + 1), 0);
    for (int64_t k : cid) c.start[k + 1]++;
    for (int64_t k = 0; k < nx * ny; ++k) c.start[k + 1] += c.start[k];
    c.items.resize(item.size());
    std::vector<int32_t> fill(c.start.begin(), c.start.end() - 1);
    for (size_t i = 0; i < item.size(); ++i) c.items[fill[cid[i]]++] = item[i];
    return c;
}

struct HostBins {
    double y0, dy;
    int nb;
    std::vector<int32_t> start, items;
};

// reference scene/scene.py:314-331 (SurfaceRep._ybins)
HostBins build_bins(const double* V, const int32_t* E, int ne) {
    HostBins b;
    std::vector<double> ylo(ne), yhi(ne);
    double ymin = INFINITY, ymax = -INFINITY;
    for (int e = 0; e < ne; ++e) {
        const double ay = V[2 * E[2 * e] + 1], by = V[2 * E[2 * e + 1] + 1];
        ylo[e] = std::min(ay, by);
        yhi[e] = std::max(ay, by);
        ymin = std::min(ymin, ylo[e]);
        ymax = std::max(ymax, yhi[e]);
    }
    b.nb = std::max(1, (int)std::floor(std::sqrt((double)ne)) * 2);
    b.y0 = ymin;
    b.dy = std::max(ymax - b.y0, 1e-300) / b.nb;
    std::vector<int64_t> bin;
    std::vector<int32_t> item;
    for (int e = 0; e < ne; ++e) {
        const int64_t i0 = std::clamp<int64_t>((int64_t)std::floor((ylo[e] - b.y0) / b.dy), 0, b.nb - 1);
        const int64_t i1 = std::clamp<int64_t>((int64_t)std::floor((yhi[e] - b.y0) / b.dy), 0, b.nb - 1);
        for (int64_t k = i0; k <= i1; ++k) {
            bin.push_back(k);
            item.push_back(e);
        }
    }
    b.start.assign((size_t)b.nb + 1, 0);
    for (int64_t k : bin) b.start[k + 1]++;
    for (int k = 0; k < b.nb; ++k) b.start[k + 1] += b.start[k];
    b.items.resize(item.size());
    std::vector<int32_t> fill(b.start.begin(), b.start.end() - 1);
    for (size_t i = 0; i < item.size(); ++i) b.items[fill[bin[i]]++] = item[i];
    return b;
}

// Bounding-volume tree over the edges for the nearest-edge search of points that no cell list reaches: node i holds
// the box (min x, min y, max x, max y) of its edges' endpoints and link (right, 0) for an inner node, whose left child
// is i + 1, or (-(first + 1), count) for a leaf of items[first .. first + count).  Median split of the edge centres on
// the longer axis of their bounds; leaves of at most kLeaf edges, so the depth is at most ceil(log2(ne / kLeaf)) + 1.
constexpr int kLeaf = 4;
constexpr int kTreeStack = 64;

struct HostTree {
    std::vector<double4> box;
    std::vector<int2> link;
    std::vector<int32_t> items;
    int depth = 0;
};

static void tree_node(HostTree& t, const double* V, const int32_t* E, const std::vector<double>& cx,
                      const std::vector<double>& cy, int first, int count, int depth) {
    const int me = (int)t.box.size();
    double4 b{INFINITY, INFINITY, -INFINITY, -INFINITY};
    double lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
    for (int i = first; i < first + count; ++i) {
        const int e = t.items[i];
        for (int j = 0; j < 2; ++j) {
            const double x = V[2 * E[2 * e + j]], y = V[2 * E[2 * e + j] + 1];
            b.x = std::min(b.x, x); b.y = std::min(b.y, y); b.z = std::max(b.z, x); b.w = std::max(b.w, y);
        }
        lx = std::min(lx, cx[e]); hx = std::max(hx, cx[e]); ly = std::min(ly, cy[e]); hy = std::max(hy, cy[e]);
    }
    t.box.push_back(b);
    t.link.push_back(int2{0, 0});
    t.depth = std::max(t.depth, depth);
    if (count <= kLeaf) {
        t.link[me] = int2{-(first + 1), count};
        return;
    }
    const bool by_x = (hx - lx) >= (hy - ly);
    const int half = count / 2;
    auto beg = t.items.begin() + first;
    std::nth_element(beg, beg + half, beg + count, [&](int32_t a, int32_t c) {
        const double ka = by_x ? cx[a] : cy[a], kc = by_x ? cx[c] : cy[c];
        return ka < kc || (ka == kc && a < c);
    });
    tree_node(t, V, E, cx, cy, first, half, depth + 1);
    t.link[me].x = (int)t.box.size();
    tree_node(t, V, E, cx, cy, first + half, count - half, depth + 1);
}

HostTree build_tree(const double* V, const int32_t* E, int ne) {
    HostTree t;
    std::vector<double> cx(ne), cy(ne);
    for (int e = 0; e < ne; ++e) {
        cx[e] = 0.5 * (V[2 * E[2 * e]] + V[2 * E[2 * e + 1]]);
        cy[e] = 0.5 * (V[2 * E[2 * e] + 1] + V[2 * E[2 * e + 1] + 1]);
    }
    t.items.resize(ne);
    for (int e = 0; e < ne; ++e) t.items[e] = e;
    t.box.reserve(2 * (size_t)ne / kLeaf + 2);
    t.link.reserve(2 * (size_t)ne / kLeaf + 2);
    tree_node(t, V, E, cx, cy, 0, ne, 1);
    return t;
}

// Reference points of the local winding count (local_winding_edge): for each cell with edges, the first of a
// few trial points inside it whose ray decisions over its y-bin (winding_point's test) and whose orientations
// against every edge of the cell's list are certain; (x, y, indicator, 1), or zeros when no trial point qualifies.
// `local` is cleared when an edge outside a point's cell list could be uncertain for the reference's ray at it:
// such an edge crosses the ray farther than smax from the point, so its line passes within 32 u D^2 / (len smax)
// of the point (D the grid diagonal) only if |dy| <= 32 u D^2 / smax; exactly horizontal edges never count.
double4 cell_reference(int k, const HostCells& c, const HostBins& b, const std::vector<double4>& cxy,
                       const std::vector<double4>& bxy, const double* V, const int32_t* E, int ne, int background) {
    static const double trial[][2] = {{0.5, 0.5},   {0.31, 0.67}, {0.69, 0.33}, {0.33, 0.31}, {0.67, 0.71},
                                      {0.5, 0.19},  {0.19, 0.5},  {0.81, 0.5},  {0.5, 0.83},  {0.41, 0.57},
                                      {0.23, 0.23}, {0.77, 0.21}, {0.24, 0.79}, {0.79, 0.77}, {0.57, 0.39}};
    const int cx = k / c.ny, cy = k % c.ny;
    for (const auto& f : trial) {
        const double qx = c.lo0 + ((double)cx + f[0]) * c.cell;
        const double qy = c.lo1 + ((double)cy + f[1]) * c.cell;
        bool ok = true;
        for (int i = c.start[k]; ok && i < c.start[k + 1]; ++i)
            ok = orient_certain(cxy[i].x, cxy[i].y, cxy[i].z, cxy[i].w, qx, qy) != 0;
        const double iyf = std::floor((qy - b.y0) / b.dy);
        if (ok && iyf >= 0.0 && iyf <= (double)b.nb) {
            const int iy = (int)iyf < b.nb - 1 ? (int)iyf : b.nb - 1;
            for (int i = b.start[iy]; ok && i < b.start[iy + 1]; ++i) {
                const double4 e = bxy[i];
                if (e.x < qx && e.z < qx) continue;
                ok = orient_certain(e.x, e.y, e.z, e.w, qx, qy) != 0;
            }
        }
        if (!ok) continue;
        const double w = winding_point(qx, qy, c.cell, V, E, ne, b.start.data(), bxy.data(), b.y0, b.dy, b.nb,
                                       (double)background);
        return double4{qx, qy, w, 1.0};
    }
    return double4{0.0, 0.0, 0.0, 0.0};
}

// ------------------------------------------------------------------------------------------------ device kernels
struct AdjArgs {
    const double* lpos;
    const double* lsup;
    const uint8_t* valid;
    const int32_t* order;
    int n;
    const double* V;
    const int32_t* E;
    int ne;
    double clo0, clo1, cell;
    int nx, ny;
    const int32_t* cstart;
    const int32_t* citems;
    int K;
    const double* cind;
    const int32_t* cempty;
    double bg, y0, dy;
    int nb;
    const int32_t* bstart;
    const double4* bxy;
    const double4* cxy;
    const double4* cref;        // nullptr: no local winding
    int32_t* slots;
    int32_t* counts;
    double* ind;
    int pad;
    double4* prow;              // non-null: the sorted-position layout (slots, counts, prow indexed by position)
};

// one thread per row (reference scene/fixedadj.py:67-147); with `order` (rows sorted by cell) thread i serves
// row order[i], so the rows of one wavefront share their cell's item list.  A candidate whose segment box lies
// farther than h (1 + 1e-9) from the point on either axis cannot pass the reference predicate (the rounded
// distance then exceeds h), so it is rejected before the division.  The indicator of a row in a cell with edges is
// the cell reference point's plus the crossings of the segment from it, counted over the same item list
// (local_winding_edge); a row with an uncertain decision, or in a cell without a reference point, takes the
// reference's ray over its y-bin (winding_point).
__global__ __launch_bounds__(256) void adjacency2d_kernel(AdjArgs a) {
    const int i_ = blockIdx.x * blockDim.x + threadIdx.x;
    if (i_ >= a.n) return;
    const int r = a.order ? a.order[i_] : i_;
    const int64_t w = a.prow ? i_ : r;          // the index the outputs are written at
    int n = 0;
    double ind = 0.0;
    const double px = a.lpos[2 * r], py = a.lpos[2 * r + 1];
    const double h = a.lsup[r];
    if (a.valid == nullptr || a.valid[r] != 0) {
        const double fx = floor((px - a.clo0) / a.cell);
        const double fy = floor((py - a.clo1) / a.cell);
        const bool inside = fx >= 0.0 && fx < (double)a.nx && fy >= 0.0 && fy < (double)a.ny;
        if (inside) {
            const int cid = (int)fx * a.ny + (int)fy;
            const double reach = h * (1.0 + 1e-9) + 1e-14 * (1.0 + fabs(px) + fabs(py));
            const bool empty = a.cempty[cid] != 0;
            double4 q = double4{0.0, 0.0, 0.0, 0.0};
            if (!empty && a.cref) q = a.cref[cid];
            const bool local = q.w != 0.0;
            int dw = 0;
            bool unsure = false;
            // four items' records and ids are requested together, then tested in list order (the loop waits on load
            // latency, not on arithmetic)
            const int i0 = a.cstart[cid], i1 = a.cstart[cid + 1];
            for (int it = i0; it < i1; it += 4) {
                double4 rec[4];
                int item[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = min(it + u, i1 - 1);
                    rec[u] = a.cxy[j];
                    item[u] = a.citems[j];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (it + u >= i1) break;
                    const double ax = rec[u].x, ay = rec[u].y;
                    const double bx = rec[u].z, by = rec[u].w;
                    if (local) local_winding_edge(ax, ay, bx, by, px, py, q.x, q.y, dw, unsure);
                    if (fmin(ax, bx) - px > reach || px - fmax(ax, bx) > reach || fmin(ay, by) - py > reach ||
                        py - fmax(ay, by) > reach)
                        continue;
                    const double dx = bx - ax;
                    const double dyy = by - ay;
                    double t = ((px - ax) * dx + (py - ay) * dyy) / fmax(dx * dx + dyy * dyy, 1.0e-300);
                    t = fmin(fmax(t, 0.0), 1.0);
                    const double ex = px - (ax + t * dx);
                    const double ey = py - (ay + t * dyy);
                    if (sqrt(ex * ex + ey * ey) < h) {
                        a.slots[w * a.K + n] = item[u];
                        ++n;
                    }
                }
            }
            if (empty) {
                ind = a.cind[cid];
            } else if (local && !unsure) {
                ind = q.z + (double)dw;
            } else {
                ind = winding_point(px, py, h, a.V, a.E, a.ne, a.bstart, a.bxy, a.y0, a.dy, a.nb, a.bg);
            }
        } else {
            ind = a.bg;
        }
    }
    if (a.pad)
        for (int s = n; s < a.K; ++s) a.slots[w * a.K + s] = -1;
    if (a.counts) a.counts[w] = n;
    if (a.prow) a.prow[w] = double4{px, py, h, ind};
    else a.ind[r] = ind;
}

struct SdArgs {
    const double* lpos;
    int m;
    double smax;
    const double* V;
    const int32_t* E;
    int ne;
    double clo0, clo1, cell;
    int nx, ny;
    const int32_t* cstart;
    const int32_t* citems;
    const double* cind;
    const int32_t* cempty;
    double bg, y0, dy;
    int nb;
    const int32_t* bstart;
    const double4* bxy;
    const double4* cxy;
    const double4* cref;        // nullptr: no local winding
    const double4* tbox;
    const int2* tlink;
    const int32_t* titems;
    double extent;
    double* d;
    double* nrm;
    int32_t* edge;
    int ng;                     // nearest-edge candidate grid (0: none)
    double nlo0, nlo1, ncs;
    const int32_t* nstart;
    const int32_t* nitems;
};

// distance from (px, py) to the segment (ax, ay) - (bx, by) and its closest point
__device__ inline double seg_dist_xy(double px, double py, double ax, double ay, double bx, double by, double* cpx,
                                     double* cpy) {
    const double ex = bx - ax, ey = by - ay;
    double t = ((px - ax) * ex + (py - ay) * ey) / (ex * ex + ey * ey);
    t = fmin(fmax(t, 0.0), 1.0);
    *cpx = ax + t * ex;
    *cpy = ay + t * ey;
    const double dx = px - *cpx, dy = py - *cpy;
    return sqrt(dx * dx + dy * dy);
}

__device__ inline double seg_dist(double px, double py, const double* V, const int32_t* E, int e, double* cpx, double* cpy) {
    const double ax = V[2 * E[2 * e]], ay = V[2 * E[2 * e] + 1];
    const double ex = V[2 * E[2 * e + 1]] - ax, ey = V[2 * E[2 * e + 1] + 1] - ay;
    double t = ((px - ax) * ex + (py - ay) * ey) / (ex * ex + ey * ey);
    t = fmin(fmax(t, 0.0), 1.0);
    *cpx = ax + t * ex;
    *cpy = ay + t * ey;
    const double dx = px - *cpx, dy = py - *cpy;
    return sqrt(dx * dx + dy * dy);
}

// distance from (px, py) to the box (x0, y0, x1, y1), 0 inside
__device__ inline double box_dist(double px, double py, double4 b) {
    const double dx = fmax(fmax(b.x - px, px - b.z), 0.0);
    const double dy = fmax(fmax(b.y - py, py - b.w), 0.0);
    return sqrt(dx * dx + dy * dy);
}

// the nearest edge over the tree (the lowest edge id among equal distances), starting from the candidate (best, k):
// a subtree is skipped only when its box is farther than best by more than the rounding of seg_dist can account for
// (slack), so every edge whose computed distance could equal or beat best is evaluated
struct TreeView {
    const double* V;
    const int32_t* E;
    const double4* tbox;
    const int2* tlink;
    const int32_t* titems;
    double extent;
};

__device__ inline double tree_slack(double px, double py, double extent) {
    return 256.0 * 1.1102230246251565e-16 * (fabs(px) + fabs(py) + extent);
}

__device__ inline void tree_nearest(const TreeView& a, double px, double py, double& best, int& k, double& bx, double& by) {
    const double slack = tree_slack(px, py, a.extent);
    int stack[kTreeStack];
    int sp = 0, node = 0;
    for (;;) {
        const int2 l = a.tlink[node];
        if (l.x < 0) {
            const int first = -l.x - 1;
            for (int i = first; i < first + l.y; ++i) {
                const int e = a.titems[i];
                double qx, qy;
                const double dd = seg_dist(px, py, a.V, a.E, e, &qx, &qy);
                if (dd < best || (dd == best && e < k)) { best = dd; k = e; bx = qx; by = qy; }
            }
        } else {
            const int lc = node + 1, rc = l.x;
            const double dl = box_dist(px, py, a.tbox[lc]), dr = box_dist(px, py, a.tbox[rc]);
            const bool gl = dl <= best + slack, gr = dr <= best + slack;
            if (gl && gr) {                 // nearer child first, the other on the stack
                stack[sp++] = dl <= dr ? rc : lc;
                node = dl <= dr ? lc : rc;
                continue;
            }
            if (gl || gr) {
                node = gl ? lc : rc;
                continue;
            }
        }
        // pop the next subtree that can still hold a candidate
        bool found = false;
        while (sp > 0) {
            const int c = stack[--sp];
            if (box_dist(px, py, a.tbox[c]) <= best + slack) { node = c; found = true; break; }
        }
        if (!found) return;
    }
}

// the edges whose computed distance from (px, py) is at most r: their number, the first `cap` of them written to out
// (tree order; out nullable)
__device__ inline int tree_within(const TreeView& a, double px, double py, double r, int32_t* out, int cap) {
    const double slack = tree_slack(px, py, a.extent);
    int stack[kTreeStack];
    int sp = 0, node = 0, n = 0;
    for (;;) {
        const int2 l = a.tlink[node];
        if (l.x < 0) {
            const int first = -l.x - 1;
            for (int i = first; i < first + l.y; ++i) {
                const int e = a.titems[i];
                double qx, qy;
                if (seg_dist(px, py, a.V, a.E, e, &qx, &qy) <= r) {
                    if (out && n < cap) out[n] = e;
                    ++n;
                }
            }
        } else {
            const int lc = node + 1, rc = l.x;
            const bool gl = box_dist(px, py, a.tbox[lc]) <= r + slack, gr = box_dist(px, py, a.tbox[rc]) <= r + slack;
            if (gl && gr) {
                stack[sp++] = rc;
                node = lc;
                continue;
            }
            if (gl || gr) {
                node = gl ? lc : rc;
                continue;
            }
        }
        if (sp == 0) return n;
        node = stack[--sp];
    }
}

// Nearest-edge candidates of one grid cell (one thread per cell): from the cell's centre c, d_c = the distance of the
// nearest edge; every point p of the cell (half-diagonal hd) has its nearest edge e* at |p - e*| <= |p - e_c| <= d_c
// + hd, so |c - e*| <= d_c + 2 hd: the edges within d_c + 2 hd (+ the rounding of the distances, four times the tree's
// slack) hold every edge that can be nearest, ties included.  Pass 1 (out == nullptr) counts them; pass 2 writes the
// lists of the cells whose count fits NC_CAP (the others keep an empty list: their points search the tree).
constexpr int NC_CAP = 256;
constexpr int NC_G = 256;
__global__ __launch_bounds__(64) void ncand_kernel(TreeView t, double lo0, double lo1, double cs, int G, int32_t* counts,
                                                   const int32_t* start, int32_t* items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G * G) return;
    const double cx = lo0 + ((double)(i / G) + 0.5) * cs, cy = lo1 + ((double)(i % G) + 0.5) * cs;
    const double hd = 0.7071067811865476 * cs * (1.0 + 1e-12);
    double best = INFINITY, bx, by;
    int k = -1;
    tree_nearest(t, cx, cy, best, k, bx, by);
    const double r = best + 2.0 * hd + 4.0 * tree_slack(fabs(cx) + hd, fabs(cy) + hd, t.extent);
    if (!items) {
        counts[i] = tree_within(t, cx, cy, r, nullptr, 0);
    } else {
        const int n = start[i + 1] - start[i];
        if (n > 0) tree_within(t, cx, cy, r, items + start[i], n);
    }
}

// nearest edge, sign from the winding number at support normalisation `smax` (reference scene/scene.py:1312-1372
// with the sync-free indicator of scene/fixedadj.py:273-276); the cell list gives the exact nearest edge whenever
// it is closer than smax, else the edge tree is searched from the cell list's candidate
__global__ __launch_bounds__(256) void signed_distance2d_kernel(SdArgs a) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.m) return;
    const double px = a.lpos[2 * r], py = a.lpos[2 * r + 1];
    const double fx = floor((px - a.clo0) / a.cell);
    const double fy = floor((py - a.clo1) / a.cell);
    const bool inside = fx >= 0.0 && fx < (double)a.nx && fy >= 0.0 && fy < (double)a.ny;
    const int cid = inside ? (int)fx * a.ny + (int)fy : 0;
    double best = INFINITY, bx = 0, by = 0;
    int k = -1;
    double ind = a.bg;
    if (inside) {
        // the cell's items (ascending edge id, the endpoints in their records): the nearest edge among them, and
        // the winding number by the local crossings from the cell's reference point (as the adjacency)
        const bool empty = a.cempty[cid] != 0;
        double4 q = double4{0.0, 0.0, 0.0, 0.0};
        if (!empty && a.cref) q = a.cref[cid];
        const bool local = q.w != 0.0;
        int dw = 0;
        bool unsure = false;
        for (int i = a.cstart[cid]; i < a.cstart[cid + 1]; ++i) {
            const double4 rec = a.cxy[i];
            if (local) local_winding_edge(rec.x, rec.y, rec.z, rec.w, px, py, q.x, q.y, dw, unsure);
            double qx, qy;
            const double dd = seg_dist_xy(px, py, rec.x, rec.y, rec.z, rec.w, &qx, &qy);
            if (dd < best) {
                best = dd;
                k = a.citems[i];
                bx = qx;
                by = qy;
            }
        }
        if (empty) ind = a.cind[cid];
        else if (local && !unsure) ind = q.z + (double)dw;
        else ind = winding_point(px, py, a.smax, a.V, a.E, a.ne, a.bstart, a.bxy, a.y0, a.dy, a.nb, a.bg);
    }
    if (!(best < a.cell)) {                 // cell = the support the list was inflated by
        // the candidates of the point's cell in the nearest-edge grid, else the edge tree
        const double gx = floor((px - a.nlo0) / a.ncs), gy = floor((py - a.nlo1) / a.ncs);
        const int G = a.ng;
        const bool in_grid = G > 0 && gx >= 0.0 && gx < (double)G && gy >= 0.0 && gy < (double)G;
        const int gc = in_grid ? (int)gx * G + (int)gy : 0;
        const int j0 = in_grid ? a.nstart[gc] : 0, j1 = in_grid ? a.nstart[gc + 1] : 0;
        if (j1 > j0) {
            for (int j = j0; j < j1; ++j) {
                const int e = a.nitems[j];
                double qx, qy;
                const double dd = seg_dist(px, py, a.V, a.E, e, &qx, &qy);
                if (dd < best || (dd == best && e < k)) {
                    best = dd;
                    k = e;
                    bx = qx;
                    by = qy;
                }
            }
        } else {
            tree_nearest(TreeView{a.V, a.E, a.tbox, a.tlink, a.titems, a.extent}, px, py, best, k, bx, by);
        }
    }
    const bool ins = ind > 0.5;
    a.d[r] = ins ? -best : best;
    double n0 = 0, n1 = 0;
    if (k >= 0) {
        const double ex = a.V[2 * a.E[2 * k + 1]] - a.V[2 * a.E[2 * k]];
        const double ey = a.V[2 * a.E[2 * k + 1] + 1] - a.V[2 * a.E[2 * k] + 1];
        const double el = sqrt(ex * ex + ey * ey);
        n0 = ey / el;
        n1 = -ex / el;
        const double vx = px - bx, vy = py - by;
        const double dn = sqrt(vx * vx + vy * vy);
        if (dn > 1e-12 && !ins) {
            n0 = vx / fmax(dn, 1e-300);
            n1 = vy / fmax(dn, 1e-300);
        }
    }
    a.nrm[2 * r] = n0;
    a.nrm[2 * r + 1] = n1;
    if (a.edge) a.edge[r] = k;
}

// One-pass counting sort by cell (grids too fine for the two-pass sort below): the cell key of every row and its
// rank among the rows of its key, then each row to its key's offset plus its rank.  A row inside the grid takes its
// rank from an atomic count of its cell; the rows outside it (often most of a workload) take the key nbins + their
// 64-lane wavefront and their rank among that wavefront's outside rows from a ballot, so they need no atomic (and
// land last, in input order).  The order of the rows inside one cell depends on the atomics' timing; it only decides
// which thread serves a row, never a result.
__global__ __launch_bounds__(256) void cell_rank_kernel(const double* lpos, int n, double clo0, double clo1, double cell, int nx,
                                                        int ny, int32_t* key, int32_t* rank, int32_t* hist) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double fx = floor((lpos[2 * r] - clo0) / cell);
    const double fy = floor((lpos[2 * r + 1] - clo1) / cell);
    const bool inside = fx >= 0.0 && fx < (double)nx && fy >= 0.0 && fy < (double)ny;
    const uint64_t out = __ballot(!inside);
    if (inside) {
        const int k = (int)fx * ny + (int)fy;
        key[r] = k;
        rank[r] = atomicAdd(hist + k, 1);
    } else {
        const int k = nx * ny + (r >> 6);
        const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(out >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)out, 0u));
        key[r] = k;
        rank[r] = below;
        if (below == 0) hist[k] = __popcll(out);
    }
}
constexpr int PLACE_RPT = 4;
__global__ __launch_bounds__(256) void cell_place_kernel(const int32_t* key, const int32_t* rank, const int32_t* start, int n,
                                                         int32_t* order) {
    const int64_t r0 = (int64_t)blockIdx.x * 256 * PLACE_RPT + threadIdx.x;
    int k[PLACE_RPT], rk[PLACE_RPT], st[PLACE_RPT];
#pragma unroll
    for (int j = 0; j < PLACE_RPT; ++j) {
        const int64_t r = r0 + 256 * j;
        k[j] = r < n ? key[r] : 0;
        rk[j] = r < n ? rank[r] : 0;
    }
#pragma unroll
    for (int j = 0; j < PLACE_RPT; ++j) st[j] = r0 + 256 * j < n ? start[k[j]] : 0;
#pragma unroll
    for (int j = 0; j < PLACE_RPT; ++j)
        if (r0 + 256 * j < n) order[st[j] + rk[j]] = (int)(r0 + 256 * j);
}
// The 2-D key pass of the two-pass cell sort (cellsort.h): the cell key of every row computed from its position
// and the chunk's LDS histogram of coarse bins in one pass.
__global__ __launch_bounds__(CS_T) void cs_hist_kernel(const double* lpos, int n, double clo0, double clo1, double cell, int nx,
                                                       int ny, int shift, int nc, int nblk, int32_t* key, int32_t* ghist) {
    __shared__ int h[CS_NC];
    for (int i = threadIdx.x; i < nc; i += CS_T) h[i] = 0;
    const int64_t c0 = (int64_t)blockIdx.x * CS_CH;
    // every position of the chunk is loaded before the first atomic, so the loads overlap
    double2 p[CS_RPT];
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        p[j] = r < n ? reinterpret_cast<const double2*>(lpos)[r] : double2{0.0, 0.0};
    }
    int k[CS_RPT];
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        const double fx = floor((p[j].x - clo0) / cell);
        const double fy = floor((p[j].y - clo1) / cell);
        k[j] = fx >= 0.0 && fx < (double)nx && fy >= 0.0 && fy < (double)ny ? (int)fx * ny + (int)fy : -1;
        if (r < n) key[r] = k[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const bool live = c0 + j * CS_T + threadIdx.x < n;
        const uint64_t out = __ballot(live && k[j] < 0);
        if (live && k[j] >= 0) atomicAdd(&h[k[j] >> shift], 1);
        else if (live && lane_rank(out) == 0) atomicAdd(&h[nc - 1], __popcll(out));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += CS_T) ghist[(int64_t)i * nblk + blockIdx.x] = h[i];
}


__global__ void zero_kernel(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// one-pass sort: key | rank | hist | start | scan temp (the two-pass sort: cellsort.h's layout)
struct SortLayout {
    size_t key, rank, hist, start, tmp, tmp_bytes, total;
};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// keys: the nbins cells, then one per 64-row wavefront for the rows outside the grid
inline int sort_keys(int n, int nbins) { return nbins + (n + 63) / 64; }

SortLayout sort_layout(int n, int nbins) {
    SortLayout L{};
    size_t tmp = 0;
    const int nk = sort_keys(n, nbins);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const int32_t*)nullptr, (int32_t*)nullptr, nk);
    L.key = 0;
    L.rank = al256(L.key + sizeof(int32_t) * (size_t)n);
    L.hist = al256(L.rank + sizeof(int32_t) * (size_t)n);
    L.start = al256(L.hist + sizeof(int32_t) * (size_t)nk);
    L.tmp = al256(L.start + sizeof(int32_t) * (size_t)nk);
    L.tmp_bytes = tmp;
    L.total = al256(L.tmp + tmp);
    return L;
}

template <class T>
int upload(T** dst, const T* src, size_t count, int64_t* bytes) {
    const size_t nb = std::max<size_t>(count, 1) * sizeof(T);
    WSB_HIP(hipMalloc(dst, nb));
    if (count) WSB_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    *bytes += (int64_t)nb;
    return WSB_OK;
}

}  // namespace

extern "C" int wsb_geom2d_create(const double* verts, int nv, const int32_t* edges, int ne, int background,
                                 double support_max, wsb_geom2d** out, int* K) {
    WSB_REQUIRE(verts && edges && out, WSB_EINVAL, "null argument");
    WSB_REQUIRE(nv >= 2 && ne >= 1, WSB_EINVAL, "geometry needs at least one edge (nv=%d ne=%d)", nv, ne);
    WSB_REQUIRE(support_max > 0 && std::isfinite(support_max), WSB_EINVAL, "support_max must be positive and finite");
    for (int i = 0; i < 2 * ne; ++i)
        WSB_REQUIRE(edges[i] >= 0 && edges[i] < nv, WSB_EINVAL, "edge vertex index %d out of range", edges[i]);
    for (int i = 0; i < 2 * nv; ++i) WSB_REQUIRE(std::isfinite(verts[i]), WSB_EINVAL, "non-finite vertex");
    HostCells c = build_cells(verts, edges, ne, support_max);
    HostBins b = build_bins(verts, edges, ne);
    HostTree tree = build_tree(verts, edges, ne);
    WSB_REQUIRE(tree.depth < kTreeStack, WSB_EINVAL, "edge tree depth %d exceeds %d", tree.depth, kTreeStack);
    auto rec_of = [&](int e) {
        return double4{verts[2 * edges[2 * e]], verts[2 * edges[2 * e] + 1], verts[2 * edges[2 * e + 1]],
                       verts[2 * edges[2 * e + 1] + 1]};
    };
    std::vector<double4> cxy(c.items.size()), bxy(b.items.size()), erec(ne), etan(ne);
    for (size_t i = 0; i < c.items.size(); ++i) cxy[i] = rec_of(c.items[i]);
    for (size_t i = 0; i < b.items.size(); ++i) bxy[i] = rec_of(b.items[i]);
    for (int e = 0; e < ne; ++e) {
        erec[e] = rec_of(e);
        const double dx = erec[e].z - erec[e].x, dy = erec[e].w - erec[e].y;
        const double len = std::sqrt(dx * dx + dy * dy);
        etan[e] = len > 0 ? double4{dx / len, dy / len, 1.0 / len, len} : double4{0, 0, 0, 0};
    }
    const int ncells = c.nx * c.ny;
    std::vector<double> cind(ncells, 0.0);
    std::vector<int32_t> cempty(ncells, 0);
    std::vector<double4> cref(ncells, double4{0.0, 0.0, 0.0, 0.0});
    int local = 1;
    int64_t noref = 0;
    {
        const double gx = c.nx * c.cell, gy = c.ny * c.cell;
        const double thr = 64.0 * 1.1102230246251565e-16 * (gx * gx + gy * gy) / support_max;
        for (int e = 0; e < ne && local; ++e) {
            const double dyv = erec[e].w - erec[e].y;
            if (dyv != 0.0 && std::fabs(dyv) <= thr) local = 0;
        }
    }
    int Kmax = 1;
    for (int k = 0; k < ncells; ++k) {
        const int cnt = c.start[k + 1] - c.start[k];
        Kmax = std::max(Kmax, cnt);
        if (cnt != 0 && local) {
            cref[k] = cell_reference(k, c, b, cxy, bxy, verts, edges, ne, background);
            noref += cref[k].w == 0.0;
        }
        if (cnt == 0) {
            cempty[k] = 1;
            const int cx = k / c.ny, cy = k % c.ny;
            const double px = c.lo0 + ((double)cx + 0.5) * c.cell;
            const double py = c.lo1 + ((double)cy + 0.5) * c.cell;
            cind[k] = winding_point(px, py, 0.0, verts, edges, ne, b.start.data(), bxy.data(), b.y0, b.dy, b.nb,
                                    (double)background);
        }
    }
    auto* g = new wsb_geom2d();
    WSB_HIP(hipGetDevice(&g->device));
    g->nv = nv;
    g->ne = ne;
    g->background = background;
    g->support_max = support_max;
    g->clo0 = c.lo0;
    g->clo1 = c.lo1;
    g->cell = c.cell;
    g->nx = c.nx;
    g->ny = c.ny;
    g->K = Kmax;
    g->y0 = b.y0;
    g->dy = b.dy;
    g->nb = b.nb;
    g->ncell_items = (int64_t)c.items.size();
    g->nbin_items = (int64_t)b.items.size();
    int st = WSB_OK;
    if (!st) st = upload(&g->verts, verts, 2 * (size_t)nv, &g->bytes);
    if (!st) st = upload(&g->edges, edges, 2 * (size_t)ne, &g->bytes);
    if (!st) st = upload(&g->cstart, c.start.data(), c.start.size(), &g->bytes);
    if (!st) st = upload(&g->citems, c.items.data(), c.items.size(), &g->bytes);
    if (!st) st = upload(&g->cind, cind.data(), cind.size(), &g->bytes);
    if (!st) st = upload(&g->cempty, cempty.data(), cempty.size(), &g->bytes);
    if (!st) st = upload(&g->bstart, b.start.data(), b.start.size(), &g->bytes);
    if (!st) st = upload(&g->bitems, b.items.data(), b.items.size(), &g->bytes);
    if (!st) st = upload(&g->cxy, cxy.data(), cxy.size(), &g->bytes);
    if (!st) st = upload(&g->bxy, bxy.data(), bxy.size(), &g->bytes);
    if (!st) st = upload(&g->erec, erec.data(), erec.size(), &g->bytes);
    if (!st) st = upload(&g->etan, etan.data(), etan.size(), &g->bytes);
    if (!st) st = upload(&g->cref, cref.data(), cref.size(), &g->bytes);
    if (!st) st = upload(&g->tbox, tree.box.data(), tree.box.size(), &g->bytes);
    if (!st) st = upload(&g->tlink, tree.link.data(), tree.link.size(), &g->bytes);
    if (!st) st = upload(&g->titems, tree.items.data(), tree.items.size(), &g->bytes);
    {
        double ext = 0.0;
        for (int i = 0; i < 2 * nv; ++i) ext = std::max(ext, std::fabs(verts[i]));
        g->extent = ext;
        g->tnodes = (int)tree.box.size();
    }
    g->local = local;
    g->noref = noref;
    if (st) {
        wsb_geom2d_destroy(g);
        return st;
    }
    *out = g;
    if (K) *K = Kmax;
    return WSB_OK;
}

extern "C" int wsb_geom2d_destroy(wsb_geom2d* g) {
    if (!g) return WSB_OK;
    for (void* p : {(void*)g->verts, (void*)g->edges, (void*)g->cstart, (void*)g->citems, (void*)g->cind,
                    (void*)g->cempty, (void*)g->bstart, (void*)g->bitems, (void*)g->cxy, (void*)g->bxy, (void*)g->erec,
                    (void*)g->etan, (void*)g->cref, (void*)g->tbox, (void*)g->tlink, (void*)g->titems,
                    (void*)g->nstart, (void*)g->nitems})
        if (p) (void)hipFree(p);
    delete g;
    return WSB_OK;
}

extern "C" int wsb_geom2d_info(const wsb_geom2d* g, int64_t* info) {
    WSB_REQUIRE(g && info, WSB_EINVAL, "null argument");
    info[0] = g->K;
    info[1] = g->nx;
    info[2] = g->ny;
    info[3] = g->nb;
    info[4] = g->ne;
    info[5] = g->nv;
    info[6] = g->ncell_items;
    info[7] = g->bytes;
    info[8] = g->local;
    info[9] = g->noref;
    return WSB_OK;
}

extern "C" int wsb_geom2d_arrays(const wsb_geom2d* g, const double** verts, const int32_t** edges) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    if (verts) *verts = g->verts;
    if (edges) *edges = g->edges;
    return WSB_OK;
}

extern "C" size_t wsb_sort2d_scratch_bytes(const wsb_geom2d* g, int n) {
    if (!g || n <= 0) return 0;
    const int nbins = g->nx * g->ny;
    return two_pass_sort(nbins) ? cs_layout(n, nbins).total : sort_layout(n, nbins).total;
}

extern "C" int wsb_sort2d(const wsb_geom2d* g, const double* lpos, int n, int32_t* order, void* scratch, size_t scratch_bytes,
                          void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(lpos && order, WSB_EINVAL, "null buffer");
    const int nbins = g->nx * g->ny;
    char* base = static_cast<char*>(scratch);
    hipStream_t s = wsb::as_stream(stream);
    if (two_pass_sort(nbins)) {
        const CsLayout C = cs_layout(n, nbins);
        WSB_REQUIRE(scratch && scratch_bytes >= C.total, WSB_ENOMEM, "sort needs %zu scratch bytes (got %zu)", C.total,
                    scratch_bytes);
        const CoarseGrid cg = coarse_grid(nbins);
        const int nblk = (n + CS_CH - 1) / CS_CH;
        hipLaunchKernelGGL(cs_hist_kernel, dim3(nblk), dim3(CS_T), 0, s, lpos, n, g->clo0, g->clo1, g->cell, g->nx, g->ny,
                           cg.shift, cg.ncb + 1, nblk, reinterpret_cast<int32_t*>(base + C.key),
                           reinterpret_cast<int32_t*>(base + C.ghist));
        WSB_HIP(hipGetLastError());
        return cs_place(C, base, n, nbins, order, s);
    }
    const SortLayout L = sort_layout(n, nbins);
    WSB_REQUIRE(scratch && scratch_bytes >= L.total, WSB_ENOMEM, "sort needs %zu scratch bytes (got %zu)", L.total, scratch_bytes);
    int32_t* key = reinterpret_cast<int32_t*>(base + L.key);
    int32_t* rank = reinterpret_cast<int32_t*>(base + L.rank);
    int32_t* hist = reinterpret_cast<int32_t*>(base + L.hist);
    int32_t* start = reinterpret_cast<int32_t*>(base + L.start);
    const int nk = sort_keys(n, nbins);
    hipLaunchKernelGGL(zero_kernel, dim3((nk + 255) / 256), dim3(256), 0, s, hist, nk);
    hipLaunchKernelGGL(cell_rank_kernel, dim3((n + 255) / 256), dim3(256), 0, s, lpos, n, g->clo0, g->clo1, g->cell, g->nx,
                       g->ny, key, rank, hist);
    WSB_HIP(hipGetLastError());
    size_t tmp = L.tmp_bytes;
    WSB_HIP(hipcub::DeviceScan::ExclusiveSum(base + L.tmp, tmp, hist, start, nk, s));
    hipLaunchKernelGGL(cell_place_kernel, dim3((n + 256 * PLACE_RPT - 1) / (256 * PLACE_RPT)), dim3(256), 0, s, key, rank,
                       start, n, order);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_adjacency2d(const wsb_geom2d* g, const double* lpos, const double* lsup, const uint8_t* valid, int n,
                               const int32_t* order, int32_t* slots, int32_t* counts, double* ind, int pad, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(lpos && lsup && slots && ind, WSB_EINVAL, "null buffer");
    AdjArgs a{lpos, lsup, valid, order, n, g->verts, g->edges, g->ne, g->clo0, g->clo1, g->cell, g->nx, g->ny,
              g->cstart, g->citems, g->K, g->cind, g->cempty, (double)g->background, g->y0, g->dy, g->nb,
              g->bstart, g->bxy, g->cxy, g->local ? g->cref : nullptr, slots, counts, ind, pad, nullptr};
    hipLaunchKernelGGL(adjacency2d_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_adjacency2d_sorted(const wsb_geom2d* g, const double* lpos, const double* lsup, const uint8_t* valid,
                                      int n, const int32_t* order, int32_t* slots, int32_t* counts, double* prow, int pad,
                                      void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(lpos && lsup && order && slots && counts && prow, WSB_EINVAL, "null buffer");
    AdjArgs a{lpos, lsup, valid, order, n, g->verts, g->edges, g->ne, g->clo0, g->clo1, g->cell, g->nx, g->ny,
              g->cstart, g->citems, g->K, g->cind, g->cempty, (double)g->background, g->y0, g->dy, g->nb,
              g->bstart, g->bxy, g->cxy, g->local ? g->cref : nullptr, slots, counts, nullptr, pad,
              reinterpret_cast<double4*>(prow)};
    hipLaunchKernelGGL(adjacency2d_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

namespace {

// ------------------------------------------------------------------------------------------------ cone area, extreme
// One thread per row over its slots (the edges within one support), FP64 (angle differences near pi/2 amplify FP32
// rounding).  The pose R = (c, -s, s, c) brings the world axis / vector into the body frame (R^T w).

__device__ inline double wrap_pi(double a) {
    const double two_pi = 6.283185307179586476925286766559;
    double r = a + 3.1415926535897932384626433832795;
    r = r - two_pi * floor(r / two_pi);
    return r - 3.1415926535897932384626433832795;
}

// the part of the sector (1/2) H^2 (h - l) that the chord of an edge at distance z removes from [l, h]:
// (1/2) H^2 (c_hi - c_lo) - (1/2) z^2 (tan(c_hi - pn) - tan(c_lo - pn)) over [c_lo, c_hi] = [l, h] cap the chord's
// angular interval (pn - be, pn + be) (the 2 pi copy nearest the midpoint); 0 when empty
// (reference scene/fused.py:165-176)
__device__ inline double chord_deficit(double l, double h, double z, double H, double be, double pn) {
    const double two_pi = 6.283185307179586476925286766559;
    const double m = 0.5 * (l + h);
    const double k = floor((m - pn) / two_pi + 0.5);
    const double c_lo = fmax(l, pn + two_pi * k - be);
    const double c_hi = fmin(h, pn + two_pi * k + be);
    if (!(c_lo < c_hi)) return 0.0;
    return 0.5 * H * H * (c_hi - c_lo) - 0.5 * z * z * (tan(c_hi - pn) - tan(c_lo - pn));
}

struct ConeArgs {
    const int32_t* slots;
    const int32_t* counts;
    int K;
    const double* ind;
    const double* lpos;
    const double* lsup;
    int n;
    const double* V;
    const int32_t* E;
    const double* vec;          // world axes / vectors [n*2]
    double c, s;                // pose cos, sin
    double half_angle;
    double* out;                // cone area: out[q] (wedge), out[n + q] (full disk); extreme: out[q] (max)
};

// area(solid cap disk(p, H) cap wedge(p, axis, half_angle)) and the full-disk area: (1/2) H^2 W indicator minus the
// chord deficits of the edges within H (reference scene/fused.py:179-240), added to out
__global__ __launch_bounds__(256) void cone_area2d_kernel(ConeArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    const double pi = 3.1415926535897932384626433832795, two_pi = 6.283185307179586476925286766559;
    const double H = a.lsup[q], px = a.lpos[2 * q], py = a.lpos[2 * q + 1];
    const double wx = a.vec[2 * q], wy = a.vec[2 * q + 1];
    const double th = atan2(-a.s * wx + a.c * wy, a.c * wx + a.s * wy);
    const double al = a.half_angle;
    const bool wedge_on = al < pi;
    double a_wedge = 0.0, a_full = 0.0;
    const int cnt = a.counts[q];
    for (int i = 0; i < cnt; ++i) {
        const int e = a.slots[(int64_t)q * a.K + i];
        if (e < 0) continue;
        const int v0 = a.E[2 * e], v1 = a.E[2 * e + 1];
        const double Ax = a.V[2 * v0] - px, Ay = a.V[2 * v0 + 1] - py;
        const double Bx = a.V[2 * v1] - px, By = a.V[2 * v1 + 1] - py;
        const double cr = Ax * By - Ay * Bx;
        const double L = sqrt((Bx - Ax) * (Bx - Ax) + (By - Ay) * (By - Ay));
        if (!(L > 0.0 && fabs(cr) >= 1.0e-300)) continue;
        const double sg = cr < 0.0 ? -1.0 : 1.0;
        const double z = fabs(cr) / L;
        if (!(z < H)) continue;
        const double pa = atan2(Ay, Ax);
        const double dphi = wrap_pi(atan2(By, Bx) - pa);
        const double lo = dphi < 0.0 ? pa + dphi : pa, hi = dphi < 0.0 ? pa : pa + dphi;
        const double dx = (Bx - Ax) / L, dy = (By - Ay) / L;
        const double t = Ax * dx + Ay * dy;
        const double pn = lo + wrap_pi(atan2(Ay - t * dy, Ax - t * dx) - lo);
        const double be = acos(z / H);
        a_full += sg * chord_deficit(lo, hi, z, H, be, pn);
        if (wedge_on) {
            const double cc = lo + wrap_pi(th - lo);
            for (int kk = -1; kk <= 1; ++kk) {
                const double l = fmax(lo, cc + two_pi * kk - al), h = fmin(hi, cc + two_pi * kk + al);
                if (l < h) a_wedge += sg * chord_deficit(l, h, z, H, be, pn);
            }
        } else {
            a_wedge += sg * chord_deficit(lo, hi, z, H, be, pn);
        }
    }
    const double idv = a.ind[q];
    const double w_wedge = wedge_on ? 2.0 * al : two_pi;
    a.out[q] += 0.5 * H * H * w_wedge * idv - a_wedge;
    a.out[a.n + q] += 0.5 * H * H * two_pi * idv - a_full;
}

// max over the solid points within H of |w_hat . x_hat|, x_hat the direction from the query (reference
// scene/fused.py:243-309): 1 where the line of w meets an edge piece inside the disk (or the query is in the solid),
// else the larger |cos| at the piece's ends; max-accumulated into out
__global__ __launch_bounds__(256) void dir_extreme2d_kernel(ConeArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n) return;
    const double H = a.lsup[q], px = a.lpos[2 * q], py = a.lpos[2 * q + 1];
    const double wx = a.vec[2 * q], wy = a.vec[2 * q + 1];
    double bx = a.c * wx + a.s * wy, by = -a.s * wx + a.c * wy;
    const double nb = sqrt(bx * bx + by * by);
    if (!(nb > 0.0)) return;
    bx /= nb;
    by /= nb;
    double best = a.ind[q] > 0.5 ? 1.0 : 0.0;
    const int cnt = a.counts[q];
    for (int i = 0; i < cnt; ++i) {
        const int e = a.slots[(int64_t)q * a.K + i];
        if (e < 0) continue;
        const int v0 = a.E[2 * e], v1 = a.E[2 * e + 1];
        const double Ax = a.V[2 * v0] - px, Ay = a.V[2 * v0 + 1] - py;
        const double Dx = a.V[2 * v1] - px - Ax, Dy = a.V[2 * v1 + 1] - py - Ay;
        const double aa = Dx * Dx + Dy * Dy;
        if (!(aa > 0.0)) continue;
        const double bq = Ax * Dx + Ay * Dy;
        const double disc = bq * bq - aa * (Ax * Ax + Ay * Ay - H * H);
        if (!(disc > 0.0)) continue;
        const double sq = sqrt(disc);
        const double t0 = fmax(0.0, (-bq - sq) / aa), t1 = fmin(1.0, (-bq + sq) / aa);
        if (!(t0 < t1)) continue;
        double f = 0.0;
        const double crd = Dx * by - Dy * bx;
        if (fabs(crd) > 0.0) {
            const double ts = -(Ax * by - Ay * bx) / crd;
            if (ts >= t0 && ts <= t1) f = 1.0;
        }
        if (f < 1.0) {
            for (int k = 0; k < 2; ++k) {
                const double t = k ? t1 : t0;
                const double rx = Ax + t * Dx, ry = Ay + t * Dy;
                const double rr = sqrt(rx * rx + ry * ry);
                if (rr > 0.0) f = fmax(f, fabs(rx * bx + ry * by) / rr);
            }
        }
        best = fmax(best, f);
    }
    a.out[q] = fmax(a.out[q], best);
}

}  // namespace

extern "C" int wsb_cone_area2d(const wsb_geom2d* g, const int32_t* slots, const int32_t* counts, int K, const double* ind,
                               const double* lpos, const double* lsup, int n, const double* axes, double pose_cos,
                               double pose_sin, double half_angle, double* out, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes n=%d K=%d", n, K);
    WSB_REQUIRE(half_angle >= 0.0, WSB_EINVAL, "negative half angle");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && counts && ind && lpos && lsup && axes && out, WSB_EINVAL, "null buffer");
    ConeArgs a{slots, counts, K, ind, lpos, lsup, n, g->verts, g->edges, axes, pose_cos, pose_sin, half_angle, out};
    hipLaunchKernelGGL(cone_area2d_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_dir_extreme2d(const wsb_geom2d* g, const int32_t* slots, const int32_t* counts, int K, const double* ind,
                                 const double* lpos, const double* lsup, int n, const double* vecs, double pose_cos,
                                 double pose_sin, double* out, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes n=%d K=%d", n, K);
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && counts && ind && lpos && lsup && vecs && out, WSB_EINVAL, "null buffer");
    ConeArgs a{slots, counts, K, ind, lpos, lsup, n, g->verts, g->edges, vecs, pose_cos, pose_sin, 0.0, out};
    hipLaunchKernelGGL(dir_extreme2d_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

// the nearest-edge candidate grid of a geometry (its first signed-distance call: allocates, launches two passes of
// ncand_kernel on the null stream and waits for them)
static int build_nearest_grid(wsb_geom2d* g) {
    std::vector<double> V(2 * (size_t)g->nv);
    WSB_HIP(hipMemcpy(V.data(), g->verts, V.size() * sizeof(double), hipMemcpyDeviceToHost));
    double lo0 = V[0], lo1 = V[1], hi0 = V[0], hi1 = V[1];
    for (int i = 0; i < g->nv; ++i) {
        lo0 = std::min(lo0, V[2 * i]);
        hi0 = std::max(hi0, V[2 * i]);
        lo1 = std::min(lo1, V[2 * i + 1]);
        hi1 = std::max(hi1, V[2 * i + 1]);
    }
    double ext = std::max(hi0 - lo0, hi1 - lo1);
    if (!(ext > 0.0)) ext = 1.0;
    const int G = NC_G;
    const double cs = 3.0 * ext / G;
    const double nlo0 = 0.5 * (lo0 + hi0) - 1.5 * ext, nlo1 = 0.5 * (lo1 + hi1) - 1.5 * ext;
    const TreeView t{g->verts, g->edges, g->tbox, g->tlink, g->titems, g->extent};
    const int ncell = G * G;
    int32_t* counts = nullptr;
    WSB_HIP(hipMalloc(&counts, sizeof(int32_t) * (size_t)ncell));
    hipLaunchKernelGGL(ncand_kernel, dim3((ncell + 63) / 64), dim3(64), 0, nullptr, t, nlo0, nlo1, cs, G, counts,
                       (const int32_t*)nullptr, (int32_t*)nullptr);
    std::vector<int32_t> cnt(ncell), start(ncell + 1, 0);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipMemcpy(cnt.data(), counts, sizeof(int32_t) * (size_t)ncell, hipMemcpyDeviceToHost);
    (void)hipFree(counts);
    WSB_HIP(err);
    for (int i = 0; i < ncell; ++i) start[i + 1] = start[i] + (cnt[i] <= NC_CAP ? cnt[i] : 0);
    int32_t* nstart = nullptr;
    int32_t* nitems = nullptr;
    WSB_HIP(hipMalloc(&nstart, sizeof(int32_t) * (size_t)(ncell + 1)));
    err = hipMalloc(&nitems, sizeof(int32_t) * (size_t)std::max(start[ncell], 1));
    if (err == hipSuccess) err = hipMemcpy(nstart, start.data(), sizeof(int32_t) * (size_t)(ncell + 1), hipMemcpyHostToDevice);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(ncand_kernel, dim3((ncell + 63) / 64), dim3(64), 0, nullptr, t, nlo0, nlo1, cs, G,
                           (int32_t*)nullptr, (const int32_t*)nstart, nitems);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipStreamSynchronize(nullptr);
    if (err != hipSuccess) {
        (void)hipFree(nstart);
        if (nitems) (void)hipFree(nitems);
        WSB_HIP(err);
    }
    g->ng = G;
    g->nlo0 = nlo0;
    g->nlo1 = nlo1;
    g->ncs = cs;
    g->nitems = nitems;
    g->nstart = nstart;
    g->bytes += (int64_t)(sizeof(int32_t) * ((size_t)ncell + 1 + (size_t)std::max(start[ncell], 1)));
    return WSB_OK;
}

extern "C" int wsb_signed_distance2d(const wsb_geom2d* g, const double* lpos, int m, double support_max, double* d,
                                     double* nrm, int32_t* edge, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(m >= 0, WSB_EINVAL, "negative point count");
    if (m == 0) return WSB_OK;
    WSB_REQUIRE(lpos && d && nrm, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(support_max > 0, WSB_EINVAL, "support_max must be positive");
    if (!g->nstart) {
        const int st = build_nearest_grid(const_cast<wsb_geom2d*>(g));
        if (st != WSB_OK) return st;
    }
    SdArgs a{lpos, m, support_max, g->verts, g->edges, g->ne, g->clo0, g->clo1, g->cell, g->nx, g->ny, g->cstart,
             g->citems, g->cind, g->cempty, (double)g->background, g->y0, g->dy, g->nb, g->bstart, g->bxy,
             g->cxy, g->local ? g->cref : nullptr, g->tbox, g->tlink, g->titems, g->extent, d, nrm, edge,
             g->ng, g->nlo0, g->nlo1, g->ncs, g->nstart, g->nitems};
    hipLaunchKernelGGL(signed_distance2d_kernel, dim3((m + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}
