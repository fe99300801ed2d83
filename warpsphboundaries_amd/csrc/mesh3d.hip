// Closed oriented triangle surfaces in 3-D: static grid build (host), face adjacency + winding number (device),
// the 22 stored moments (device tiles), and host evaluations of the same arithmetic.
//
// Adjacency contract: a face is a neighbour of a particle iff the reference's whole-element test keeps it,
//     sum_a (max(0, lo_a - x_a, x_a - hi_a) / h)^2 <= 1 + 1.4210854715202004e-14      (axis order, no FMA)
// over the face's vertex box (fluxbridge/device_search.py:80-85 with each triangle an element,
// fluxbridge/geometry.py:62), listed per particle in ascending face order (the sorted (query, element) order of
// device_search.py:186-189).  The static grid only narrows the candidates: a cell lists every face whose box,
// inflated by support_max (1 + 1e-9), overlaps it.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "cellsort.h"
#include "mesh3d_device.h"
#include "wsb_internal.h"

struct wsb_mesh3d {
    int device = 0;
    bool host_only = false, soup = false;
    int nv = 0, nf = 0;
    double support_max = 0;
    m3::Grid3 hg{};                      // host view (pointers into the vectors below)
    m3::Grid3 dg{};                      // device view
    int K = 0, ncomp = 0;
    int64_t nitems = 0, ndirty = 0, bytes = 0;
    std::vector<double> frec, fbox, cchi;
    std::vector<int32_t> cstart, citems, cwalk, wdir, wstart, witems;
    std::vector<uint32_t> cwork;
    std::vector<float4> crec, wrec;
    // bounding-volume tree of the faces (wsb_adjacency3d_bvh): per node two float4, (lo.xyz, left) and (hi.xyz,
    // right) for an inner node, (lo.xyz, -(first + 1)) and (hi.xyz, count) for a leaf of items [first, first + count);
    // per item the face's FP32 box rounded outward, (lo.xyz, id) and (hi.xyz, 0), in Morton order of the face centres
    std::vector<float4> bnode, bitem;
    int bvh_depth = 0;
    float4* d_bnode = nullptr;
    float4* d_bitem = nullptr;
    double* d_frec = nullptr;
    double* d_fbox = nullptr;
    double* d_cchi = nullptr;
    int32_t* d_cstart = nullptr;
    int32_t* d_citems = nullptr;
    float4* d_crec = nullptr;
    float4* d_wrec = nullptr;
    int32_t* d_cwalk = nullptr;
    uint32_t* d_cwork = nullptr;
    int32_t* d_wdir = nullptr;
    int32_t* d_wstart = nullptr;
    int32_t* d_witems = nullptr;
};

namespace {

// whole-element test of the reference, operation order kept (no contraction)
// keep_face decided without its divisions wherever the outcome is certain: the squared distance to the exact box
// below h^2 (1 - 1e-9) keeps (the reference's rounded sum differs from the exact quotient by less than 1e-15
// relative, its threshold is 1 + 1.42e-14); only the band above takes the reference's arithmetic.  (Candidates far
// from the particle are rejected before, in FP32, by prefilter32.)
__host__ __device__ inline bool keep_face(const double* __restrict__ box, double px, double py, double pz, double h);

__host__ __device__ inline bool keep_face_fast(const double* __restrict__ box, double px, double py, double pz, double h) {
    const double h2 = h * h;
    const double dx = fmax(0.0, fmax(box[0] - px, px - box[3]));
    const double dy = fmax(0.0, fmax(box[1] - py, py - box[4]));
    const double dz = fmax(0.0, fmax(box[2] - pz, pz - box[5]));
    if (dx * dx + dy * dy + dz * dz < h2 * (1.0 - 1e-9)) return true;
    return keep_face(box, px, py, pz, h);
}

__host__ __device__ inline bool keep_face(const double* __restrict__ box, double px, double py, double pz, double h) {
#pragma clang fp contract(off)
    const double p[3] = {px, py, pz};
    double sq = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double delta = fmax(0.0, fmax(box[a] - p[a], p[a] - box[3 + a])) / h;
        sq = sq + delta * delta;
    }
    return sq <= 1.0 + 1.4210854715202004e-14;
}

// a box farther than h (1 + 1e-9) on some axis fails keep_face (its rounded normalised distance exceeds 1 + 1e-9)
__host__ __device__ inline bool box_far(const double* __restrict__ box, double px, double py, double pz, double reach) {
    return box[0] - px > reach || px - box[3] > reach || box[1] - py > reach || py - box[4] > reach || box[2] - pz > reach ||
           pz - box[5] > reach;
}

__host__ __device__ inline double exact_winding(const double* frec, int nf, double px, double py, double pz, double ih) {
    double s = 0.0;
    for (int f = 0; f < nf; ++f) s += m3::face_solid_angle<m3::FP64>(frec + m3::FREC * f, px, py, pz, ih);
    return s / (4.0 * m3::PI);
}

// ------------------------------------------------------------------------------------------------ host build
int build_host(wsb_mesh3d* m, const double* V, int nv, const int32_t* Fc, int nf, double smax) {
    const int64_t nf64 = nf;
    // FP32 rounding toward -inf / +inf (conservative boxes of the pre-filters)
    auto down = [](double v) { float f = (float)v; if ((double)f > v) f = std::nextafter(f, -INFINITY); return f; };
    auto up = [](double v) { float f = (float)v; if ((double)f < v) f = std::nextafter(f, INFINITY); return f; };
    m->frec.assign((size_t)nf64 * m3::FREC, 0.0);
    m->fbox.assign((size_t)nf64 * 6, 0.0);
    double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int f = 0; f < nf; ++f) {
        const int32_t* ix = Fc + 3 * (int64_t)f;
        double* fr = m->frec.data() + (size_t)m3::FREC * f;
        double* bx = m->fbox.data() + (size_t)6 * f;
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) fr[m3::FR_V + 3 * k + a] = V[3 * (int64_t)ix[k] + a];
        const double* v0 = fr + m3::FR_V;
        const double* v1 = fr + m3::FR_V + 3;
        const double* v2 = fr + m3::FR_V + 6;
        const double d1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const double d2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        const double N[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2], d1[0] * d2[1] - d1[1] * d2[0]};
        const double nl = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
        WSB_REQUIRE(nl > 0.0 && std::isfinite(nl), WSB_EINVAL, "degenerate triangle %d", f);
        const double n[3] = {N[0] / nl, N[1] / nl, N[2] / nl};
        for (int a = 0; a < 3; ++a) fr[m3::FR_N + a] = n[a];
        fr[m3::FR_NLEN] = nl;
        double lmax2 = 0.0;
        for (int e = 0; e < 3; ++e) {
            const double* p = fr + m3::FR_V + 3 * e;
            const double* q = fr + m3::FR_V + 3 * ((e + 1) % 3);
            lmax2 = std::max(lmax2, (q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]));
        }
        fr[m3::FR_SLIVER] = nl <= 1e-2 * lmax2 ? 1.0 : 0.0;      // 2 area <= 1e-2 longest^2
        for (int e = 0; e < 3; ++e) {
            const double* p = fr + m3::FR_V + 3 * e;
            const double* q = fr + m3::FR_V + 3 * ((e + 1) % 3);
            const double D[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
            const double len = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
            const double t[3] = {D[0] / len, D[1] / len, D[2] / len};
            const double mu[3] = {t[1] * n[2] - t[2] * n[1], t[2] * n[0] - t[0] * n[2], t[0] * n[1] - t[1] * n[0]};
            for (int a = 0; a < 3; ++a) {
                fr[m3::FR_T + 3 * e + a] = t[a];
                fr[m3::FR_MU + 3 * e + a] = mu[a];
            }
            fr[m3::FR_LEN + e] = len;
        }
        for (int a = 0; a < 3; ++a) {
            bx[a] = std::min({v0[a], v1[a], v2[a]});
            bx[3 + a] = std::max({v0[a], v1[a], v2[a]});
            glo[a] = std::min(glo[a], bx[a]);
            ghi[a] = std::max(ghi[a], bx[3 + a]);
        }
    }
    // grid: cells of support_max over the inflated mesh box, coarsened if it would exceed 2^24 cells
    const double pad = smax * (1.0 + 1e-9) + 1e-12 * (1.0 + std::max({std::fabs(glo[0]), std::fabs(glo[1]), std::fabs(glo[2]),
                                                                            std::fabs(ghi[0]), std::fabs(ghi[1]), std::fabs(ghi[2])}));
    double cell = smax;
    double ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = (ghi[a] - glo[a]) + 2.0 * pad;
    auto ncell_of = [&](double c) {
        double tot = 1.0;
        for (int a = 0; a < 3; ++a) tot *= std::max(1.0, std::ceil(ext[a] / c));
        return tot;
    };
    while (ncell_of(cell) > 16777216.0) cell *= 1.25;
    m3::Grid3& g = m->hg;
    g.cell = cell;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = glo[a] - pad;
        g.n[a] = (int)std::max(1.0, std::ceil(ext[a] / cell));
    }
    const int64_t ncells = (int64_t)g.n[0] * g.n[1] * g.n[2];
    auto clampi = [](double v, int hi) { return (int)std::min<double>(std::max<double>(v, 0.0), (double)hi); };
    // dilated candidate lists (counting sort by cell, faces ascending inside a cell) and the dirty flags
    std::vector<int32_t> cnt((size_t)ncells + 1, 0);
    std::vector<uint8_t> dirty((size_t)ncells, 0);
    const double reach = smax * (1.0 + 1e-9);
    auto range = [&](const double* bx, double r, int lo3[3], int hi3[3]) {
        for (int a = 0; a < 3; ++a) {
            const double eps = 1e-12 * (1.0 + std::fabs(bx[a]) + std::fabs(bx[3 + a]));
            lo3[a] = clampi(std::floor((bx[a] - r - eps - g.lo[a]) / cell), g.n[a] - 1);
            hi3[a] = clampi(std::floor((bx[3 + a] + r + eps - g.lo[a]) / cell), g.n[a] - 1);
        }
    };
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int32_t> fill;
        if (pass == 1) {
            for (int64_t k = 0; k < ncells; ++k) cnt[k + 1] += cnt[k];
            m->citems.assign((size_t)cnt[ncells], 0);
            fill.assign(cnt.begin(), cnt.end() - 1);
        }
        for (int f = 0; f < nf; ++f) {
            const double* bx = m->fbox.data() + (size_t)6 * f;
            int lo3[3], hi3[3];
            range(bx, reach, lo3, hi3);
            for (int i = lo3[0]; i <= hi3[0]; ++i)
                for (int j = lo3[1]; j <= hi3[1]; ++j)
                    for (int k = lo3[2]; k <= hi3[2]; ++k) {
                        const int64_t c = m3::cell_id(g, i, j, k);
                        if (pass == 0) cnt[c + 1]++;
                        else m->citems[fill[c]++] = f;
                    }
            if (pass == 0) {
                range(bx, 0.0, lo3, hi3);
                for (int i = lo3[0]; i <= hi3[0]; ++i)
                    for (int j = lo3[1]; j <= hi3[1]; ++j)
                        for (int k = lo3[2]; k <= hi3[2]; ++k) dirty[m3::cell_id(g, i, j, k)] = 1;
            }
        }
    }
    m->cstart.assign(cnt.begin(), cnt.end());
    m->nitems = (int64_t)m->citems.size();
    // per-item FP32 boxes rounded outward: a conservative pre-filter read with one load per candidate
    m->crec.resize(2 * m->citems.size());
    for (size_t it = 0; it < m->citems.size(); ++it) {
        const int f = m->citems[it];
        const double* bx = m->fbox.data() + (size_t)6 * f;
        float id;
        std::memcpy(&id, &f, sizeof(float));
        m->crec[2 * it] = float4{down(bx[0]), down(bx[1]), down(bx[2]), id};
        m->crec[2 * it + 1] = float4{up(bx[3]), up(bx[4]), up(bx[5]), 0.0f};
    }
    int K = 0;
    for (int64_t c = 0; c < ncells; ++c) K = std::max(K, m->cstart[c + 1] - m->cstart[c]);
    m->K = K;
    m->cchi.assign((size_t)ncells, 0.0);
    m->cwalk.assign((size_t)ncells, -1);
    if (!m->soup) {
        // winding number of each connected component of face-free cells, from one exact sum per component
        std::vector<int32_t> comp((size_t)ncells, -1);
        int ncomp = 0;
        std::vector<int64_t> queue;
        for (int64_t c0 = 0; c0 < ncells; ++c0) {
            if (dirty[c0] || comp[c0] >= 0) continue;
            const int id = ncomp++;
            comp[c0] = id;
            queue.assign(1, c0);
            for (size_t qi = 0; qi < queue.size(); ++qi) {
                const int64_t c = queue[qi];
                const int k = (int)(c % g.n[2]), j = (int)((c / g.n[2]) % g.n[1]), i = (int)(c / ((int64_t)g.n[1] * g.n[2]));
                const int nb[6][3] = {{i + 1, j, k}, {i - 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k + 1}, {i, j, k - 1}};
                for (auto& q : nb) {
                    if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= g.n[0] || q[1] >= g.n[1] || q[2] >= g.n[2]) continue;
                    const int64_t cq = m3::cell_id(g, q[0], q[1], q[2]);
                    if (dirty[cq] || comp[cq] >= 0) continue;
                    comp[cq] = id;
                    queue.push_back(cq);
                }
            }
            const int k = (int)(c0 % g.n[2]), j = (int)((c0 / g.n[2]) % g.n[1]), i = (int)(c0 / ((int64_t)g.n[1] * g.n[2]));
            const double px = g.lo[0] + (i + 0.5) * cell, py = g.lo[1] + (j + 0.5) * cell, pz = g.lo[2] + (k + 0.5) * cell;
            const double w = exact_winding(m->frec.data(), nf, px, py, pz, 1.0);
            const double wr = std::nearbyint(w);
            WSB_REQUIRE(std::fabs(w - wr) < 1e-6, WSB_EINVAL, "winding number %.17g of a face-free cell is not an integer", w);
            for (int64_t c : queue) m->cchi[c] = wr;
        }
        m->ncomp = ncomp;
        // dirty cells: shortest axis walk to a face-free cell or out of the grid (6 directional sweeps)
        std::vector<int32_t> best((size_t)ncells, 0x7fffffff), bdir((size_t)ncells, 0);
        for (int dir = 0; dir < 6; ++dir) {
            const int ax = dir >> 1, s = (dir & 1) ? -1 : 1;
            const int au = (ax + 1) % 3, av = (ax + 2) % 3;
            for (int u = 0; u < g.n[au]; ++u)
                for (int v = 0; v < g.n[av]; ++v) {
                    int run = 0;
                    for (int t = 0; t < g.n[ax]; ++t) {
                        const int pos = s > 0 ? g.n[ax] - 1 - t : t;
                        int ijk[3];
                        ijk[ax] = pos;
                        ijk[au] = u;
                        ijk[av] = v;
                        const int64_t c = m3::cell_id(g, ijk[0], ijk[1], ijk[2]);
                        if (!dirty[c]) {
                            run = 0;
                            continue;
                        }
                        ++run;
                        if (run < best[c]) {
                            best[c] = run;
                            bdir[c] = dir;
                        }
                    }
                }
        }
        // per cell with faces: its walk and the faces that overlap the walk's column (cross-section of the cell,
        // walked range along the axis), deduplicated and ascending
        int64_t ndirty = 0;
        m->wstart.assign(1, 0);
        std::vector<int32_t> buf;
        for (int64_t c = 0; c < ncells; ++c) {
            if (!dirty[c]) continue;
            const int d = (int)ndirty++;
            m->cwalk[c] = d;
            const int dir = bdir[c], steps = best[c];
            m->wdir.push_back(dir | (steps << 3));
            const int ax = dir >> 1, sg = (dir & 1) ? -1 : 1;
            const int k = (int)(c % g.n[2]), j = (int)((c / g.n[2]) % g.n[1]), i = (int)(c / ((int64_t)g.n[1] * g.n[2]));
            const int ijk[3] = {i, j, k};
            double lo3[3], hi3[3];
            for (int a = 0; a < 3; ++a) {
                lo3[a] = g.lo[a] + ijk[a] * cell;
                hi3[a] = g.lo[a] + (ijk[a] + 1) * cell;
            }
            const int jend = ijk[ax] + sg * (steps - 1);
            lo3[ax] = g.lo[ax] + std::min(ijk[ax], jend) * cell;
            hi3[ax] = g.lo[ax] + (std::max(ijk[ax], jend) + 1) * cell;
            buf.clear();
            for (int st = 0; st < steps; ++st) {
                int q[3] = {i, j, k};
                q[ax] += sg * st;
                const int64_t cq = m3::cell_id(g, q[0], q[1], q[2]);
                for (int it = m->cstart[cq]; it < m->cstart[cq + 1]; ++it) {
                    const int f = m->citems[it];
                    const double* bx = m->fbox.data() + (size_t)6 * f;
                    bool hit = true;
                    for (int a = 0; a < 3; ++a) {
                        const double eps = 1e-9 * cell + 1e-12 * (std::fabs(lo3[a]) + std::fabs(hi3[a]));
                        if (bx[3 + a] < lo3[a] - eps || bx[a] > hi3[a] + eps) hit = false;
                    }
                    if (hit) buf.push_back(f);
                }
            }
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            m->witems.insert(m->witems.end(), buf.begin(), buf.end());
            m->wstart.push_back((int32_t)m->witems.size());
            const int au = (ax + 1) % 3, av = (ax + 2) % 3;
            for (int f : buf) {
                const double* bx = m->fbox.data() + (size_t)6 * f;
                float id;
                std::memcpy(&id, &f, sizeof(float));
                m->wrec.push_back(float4{down(bx[au]), up(bx[3 + au]), down(bx[av]), up(bx[3 + av])});
                m->wrec.push_back(float4{down(bx[ax]), up(bx[3 + ax]), id, 0.0f});
            }
        }
        m->ndirty = ndirty;
    }
    g.cstart = m->cstart.data();
    g.citems = m->citems.data();
    g.crec = m->crec.data();
    g.cchi = m->cchi.data();
    g.cwalk = m->cwalk.data();
    g.wdir = m->wdir.data();
    g.wstart = m->wstart.data();
    g.witems = m->witems.data();
    g.wrec = m->wrec.data();
    g.frec = m->frec.data();
    g.fbox = m->fbox.data();
    g.nf = nf;
    return WSB_OK;
}

int validate_closed(const double* V, int nv, const int32_t* Fc, int nf) {
    // every undirected edge twice, in opposite directions (mesh3d/reference.py:238-245)
    std::map<std::pair<int, int>, std::pair<int, int>> edges;     // (min, max) -> (count, direction sum)
    for (int f = 0; f < nf; ++f) {
        const int32_t* ix = Fc + 3 * (int64_t)f;
        for (int e = 0; e < 3; ++e) {
            const int a = ix[e], b = ix[(e + 1) % 3];
            auto& v = edges[{std::min(a, b), std::max(a, b)}];
            v.first++;
            v.second += a < b ? 1 : -1;
        }
    }
    for (auto& kv : edges)
        WSB_REQUIRE(kv.second.first == 2 && kv.second.second == 0, WSB_EINVAL,
                    "surface must be watertight and consistently oriented (edge %d-%d)", kv.first.first, kv.first.second);
    // positive signed volume, relative to the first vertex (mesh3d/reference.py:246-250)
    double s6 = 0.0;
    const double* o = V;
    for (int f = 0; f < nf; ++f) {
        const int32_t* ix = Fc + 3 * (int64_t)f;
        double r[3][3];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) r[k][a] = V[3 * (int64_t)ix[k] + a] - o[a];
        s6 += r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) + r[0][1] * (r[1][2] * r[2][0] - r[1][0] * r[2][2]) +
              r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
    }
    WSB_REQUIRE(s6 > 0.0, WSB_EINVAL, "surface must have positive signed volume");
    (void)nv;
    return WSB_OK;
}

template <class T>
int upload(T** dst, const std::vector<T>& src, int64_t* bytes) {
    const size_t nb = std::max<size_t>(src.size(), 1) * sizeof(T);
    WSB_HIP(hipMalloc(dst, nb));
    if (!src.empty()) WSB_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    *bytes += (int64_t)nb;
    return WSB_OK;
}

// ------------------------------------------------------------------------------------------------ face tree
// Morton order of the face-box centres (10 bits per axis over the centres' box), then a binary tree whose split of
// a range is the first position where the highest bit in which its first and last codes differ is set (a range of
// equal codes is split in half), leaves of at most BVH_LEAF faces; node boxes are unions of the faces' FP32 boxes
// rounded outward, so a node whose box the FP32 filter rejects holds no face the element test keeps.
constexpr int BVH_LEAF = 4;
constexpr int BVH_STACK = 64;

static uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

struct BvhBuild {
    const std::vector<uint32_t>& code;
    const std::vector<float4>& item;
    std::vector<float4>& node;
    int depth = 0;

    int build(int first, int last, int d) {           // items [first, last]
        const int me = (int)(node.size() / 2);
        node.push_back(float4{});
        node.push_back(float4{});
        depth = std::max(depth, d);
        float4 lo{INFINITY, INFINITY, INFINITY, 0.0f}, hi{-INFINITY, -INFINITY, -INFINITY, 0.0f};
        auto grow = [&](const float4& l, const float4& u) {
            lo.x = std::min(lo.x, l.x); lo.y = std::min(lo.y, l.y); lo.z = std::min(lo.z, l.z);
            hi.x = std::max(hi.x, u.x); hi.y = std::max(hi.y, u.y); hi.z = std::max(hi.z, u.z);
        };
        if (last - first + 1 <= BVH_LEAF) {
            for (int i = first; i <= last; ++i) grow(item[2 * i], item[2 * i + 1]);
            const int tag = -(first + 1), cnt = last - first + 1;
            std::memcpy(&lo.w, &tag, sizeof(float));
            std::memcpy(&hi.w, &cnt, sizeof(float));
        } else {
            int split = (first + last) / 2;
            const uint32_t a = code[first], b = code[last];
            if (a != b) {
                const int top = 31 - __builtin_clz(a ^ b);
                int lo_i = first, hi_i = last;                   // first index whose code has bit `top` set
                while (lo_i < hi_i) {
                    const int mid = (lo_i + hi_i) / 2;
                    if ((code[mid] >> top) & 1u) hi_i = mid;
                    else lo_i = mid + 1;
                }
                split = lo_i - 1;
            }
            const int l = build(first, split, d + 1);
            const int r = build(split + 1, last, d + 1);
            grow(node[2 * l], node[2 * l + 1]);
            grow(node[2 * r], node[2 * r + 1]);
            std::memcpy(&lo.w, &l, sizeof(float));
            std::memcpy(&hi.w, &r, sizeof(float));
        }
        node[2 * me] = lo;
        node[2 * me + 1] = hi;
        return me;
    }
};

int build_bvh(wsb_mesh3d* m) {
    const int nf = m->nf;
    auto down = [](double v) { float f = (float)v; if ((double)f > v) f = std::nextafter(f, -INFINITY); return f; };
    auto up = [](double v) { float f = (float)v; if ((double)f < v) f = std::nextafter(f, INFINITY); return f; };
    std::vector<double> cx(3 * (size_t)nf);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int f = 0; f < nf; ++f)
        for (int a = 0; a < 3; ++a) {
            const double c = 0.5 * (m->fbox[6 * (size_t)f + a] + m->fbox[6 * (size_t)f + 3 + a]);
            cx[3 * (size_t)f + a] = c;
            lo[a] = std::min(lo[a], c);
            hi[a] = std::max(hi[a], c);
        }
    std::vector<uint32_t> key(nf);
    for (int f = 0; f < nf; ++f) {
        uint32_t k = 0;
        for (int a = 0; a < 3; ++a) {
            const double ext = hi[a] - lo[a];
            const double u = ext > 0 ? (cx[3 * (size_t)f + a] - lo[a]) / ext : 0.0;
            k |= spread10((uint32_t)std::min(1023.0, std::max(0.0, std::floor(u * 1024.0)))) << (2 - a);
        }
        key[f] = k;
    }
    std::vector<int32_t> ord(nf);
    for (int f = 0; f < nf; ++f) ord[f] = f;
    std::sort(ord.begin(), ord.end(), [&](int32_t p, int32_t q) { return key[p] < key[q] || (key[p] == key[q] && p < q); });
    std::vector<uint32_t> code(nf);
    m->bitem.assign(2 * (size_t)nf, float4{});
    for (int i = 0; i < nf; ++i) {
        const int f = ord[i];
        code[i] = key[f];
        const double* bx = m->fbox.data() + 6 * (size_t)f;
        float id;
        std::memcpy(&id, &f, sizeof(float));
        m->bitem[2 * (size_t)i] = float4{down(bx[0]), down(bx[1]), down(bx[2]), id};
        m->bitem[2 * (size_t)i + 1] = float4{up(bx[3]), up(bx[4]), up(bx[5]), 0.0f};
    }
    m->bnode.clear();
    m->bnode.reserve(4 * (size_t)nf / BVH_LEAF + 4);
    BvhBuild b{code, m->bitem, m->bnode};
    b.build(0, nf - 1, 1);
    m->bvh_depth = b.depth;
    WSB_REQUIRE(b.depth < BVH_STACK, WSB_EINVAL, "face tree depth %d exceeds %d", b.depth, BVH_STACK);
    return WSB_OK;
}

// ------------------------------------------------------------------------------------------------ device kernels
struct Adj3Args {
    m3::Grid3 g;
    const double* pos;
    const double* sup;
    const int32_t* order;       // nullable: thread i serves particle order[i] (particles grouped by cell)
    int n;
    int64_t* counts;            // count pass: per particle
    const int64_t* offsets;     // fill pass: CSR offsets
    int32_t* faces;
    int64_t capacity;
    double* chi;
    int32_t* wide_count;        // scratch: the number of wide particles, then their indices
    int32_t* wide_list;
    int32_t* walk_count;        // scratch: the number of particles with a long axis walk, then their indices
    int32_t* walk_list;
    int wide_min;               // candidate lists / axis walks longer than this go to the wavefront kernels
    int32_t* exact_count;       // scratch: the number of particles whose walk hit a rounding bound, then their indices
    int32_t* exact_list;
    bool no_faces;              // the face lists come from the face tree: adj3_kernel<0> evaluates the winding only
    const float4* bnode;        // the face tree (wsb_adjacency3d_bvh)
    const float4* bitem;
    int32_t* long_count;        // scratch: the rows of more than BVH_SHORT faces, then their indices
    int32_t* long_list;
    double4* prow;              // non-null: the sorted-position layout (counts, offsets, faces, chi and the particle's
                                // (x, y, z, h) at its position in order; lists hold positions)
    double hlim;                // a support above this (or NaN) sets *bad: the mesh's cells reach support_max only
    int32_t* bad;
};

// the particle served at position p of a list (a position in the sorted-position layout, else the particle itself)
__device__ __forceinline__ int adj3_row(const Adj3Args& a, int p) { return a.prow ? a.order[p] : p; }

constexpr int ADJ_UNROLL = 8;

// A particle whose cell lists more than WIDE3 candidates is scanned by one wavefront (adj3_wide_kernel) instead of one
// lane: lane k takes candidates k, k + 64, ... in chunks of 64, and the kept candidates of a chunk are placed in list
// order by a ballot and the number of kept lanes below.  Scenes whose mesh is fine against the particle spacing have
// thousands of candidates per particle and too few particles to fill the chip at one lane each.  Either way a
// particle's face list is the same ascending list.  Likewise a particle whose cell's axis walk has more than WIDE3
// items takes its winding number from adj3_walk_kernel, one wavefront per particle (the crossings are integers, so
// their sum does not depend on the order).  Launches of WIDE3_MAX_N particles or more fill the chip at one lane per
// particle; there the wavefront kernels take the particles of more than WIDE3_LARGE candidates (their 32-byte
// candidate records then load as contiguous runs instead of one cache line per lane; 0.323 -> 0.316 ms on the sphere
// cavity, 16 and 64 within 0.004 ms of 32).
constexpr int WIDE3 = 128;
constexpr int WIDE3_MAX_N = 1 << 18;
constexpr int WIDE3_LARGE = 32;

// Conservative FP32 rejection of a candidate from its FP32 box rounded outward (l, u): farther than h (1 + 1e-6) on one
// axis, or squared distance beyond h^2 (1 + 1e-5).  `tol` bounds the FP32 rounding of the particle's coordinates and of the
// differences (2^-23 (|p| + |box|) at most; tol = 1e-6 (1 + the largest coordinate) covers it with a wide margin),
// and every distance is shortened by tol before it is compared, so a rejected candidate is farther than h on that axis
// (or than h (1 + 4e-6) in all) in exact arithmetic: the element test would reject it too.
struct Near32 {
    float px, py, pz, reach, h2, tol;
};
__device__ inline Near32 near32(double px, double py, double pz, double h) {
    Near32 q;
    q.px = (float)px;
    q.py = (float)py;
    q.pz = (float)pz;
    q.tol = 1e-6f * (1.0f + fmaxf(fabsf(q.px), fmaxf(fabsf(q.py), fabsf(q.pz))));
    q.reach = (float)(h * (1.0 + 1e-6));     // above h after the FP32 rounding (relative 6e-8)
    q.h2 = (float)(h * h);
    return q;
}
// 0: rejected, 1: kept, 2: undecided (the FP64 test decides).  Kept: the distances to the box shrunk by tol (inside the
// exact box, whose bounds lie within an FP32 ulp of l and u) lengthened by tol again bound the exact distances from
// above; below h^2 (1 - 1e-5) in all, the element test keeps the candidate.
__device__ inline int classify32(const Near32& q, const float4& l, const float4& u) {
    const float ax = fmaxf(l.x - q.px, q.px - u.x), ay = fmaxf(l.y - q.py, q.py - u.y), az = fmaxf(l.z - q.pz, q.pz - u.z);
    const float dx = fmaxf(0.0f, ax - q.tol), dy = fmaxf(0.0f, ay - q.tol), dz = fmaxf(0.0f, az - q.tol);
    if (dx > q.reach || dy > q.reach || dz > q.reach) return 0;
    if (dx * dx + dy * dy + dz * dz > q.h2 * (1.0f + 1e-5f)) return 0;
    const float ux = fmaxf(0.0f, ax + 3.0f * q.tol), uy = fmaxf(0.0f, ay + 3.0f * q.tol), uz = fmaxf(0.0f, az + 3.0f * q.tol);
    return ux * ux + uy * uy + uz * uz < q.h2 * (1.0f - 1e-5f) ? 1 : 2;
}

// the reference's per-face solid-angle sum at one point, evaluated by the whole wavefront (face f on lane f % 64, a
// butterfly over the lane sums: every lane ends with the same value)
__device__ inline double wave_exact_winding(const m3::Grid3& g, double qx, double qy, double qz, double qi) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int f = lane; f < g.nf; f += 64) s += m3::face_solid_angle<m3::FP64>(g.frec + m3::FREC * (int64_t)f, qx, qy, qz, qi);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    return s / (4.0 * m3::PI);
}

// one thread per particle: candidates of its cell through the element test; the count pass also evaluates the
// winding number (face-free cell, axis walk, or -- when a predicate of the walk is within rounding -- the exact
// per-face sum, evaluated by adj3_exact_kernel for the particles listed here)
// MODE 0: pair counts, 1: face lists, 2: winding numbers by the walk (apart from the candidate loop: the walk needs
// twice the registers of the loop; the exact sum, in a kernel of its own, four times)
template <int MODE>
__global__ __launch_bounds__(256) void adj3_kernel(Adj3Args a) {
    constexpr bool FILL = MODE == 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < a.n;
    // (the later passes of the sorted layout read everything at the position: no index through the order)
    const int i = live ? ((a.order && !(MODE != 0 && a.prow)) ? a.order[t] : t) : 0;
    const int wi = a.prow ? t : i;      // where the particle's adjacency is written (its position, sorted layout)
    // MODE 2 serves only the particles the count pass left undecided (NaN winding number); MODE 1 only those with
    // faces
    double chi = 0.0;
    bool todo = live;
    if (MODE == 2 && live) {
        chi = a.chi[wi];
        todo = chi != chi;
    }
    if (FILL && live) todo = a.offsets[wi + 1] > a.offsets[wi];
    double px = 0.0, py = 0.0, pz = 0.0, h = 1.0;
    if (todo) {
        if (MODE != 0 && a.prow) {
            // the later passes of the sorted layout read the record the count pass wrote at the position
            const double4 pr = a.prow[wi];
            px = pr.x;
            py = pr.y;
            pz = pr.z;
            h = pr.w;
        } else {
            px = a.pos[3 * (int64_t)i];
            py = a.pos[3 * (int64_t)i + 1];
            pz = a.pos[3 * (int64_t)i + 2];
            h = a.sup[i];
            if (MODE == 0 && !(h <= a.hlim)) atomicOr(a.bad, 1);
        }
    }
    int c[3] = {0, 0, 0};
    const bool in = todo && m3::cell_of(a.g, px, py, pz, c);
    int64_t cnt = 0;
    int it0 = 0, it1 = 0;
    // a cell whose bit is clear in the work bitmap (small, cached) lists no candidate and has no walk: its particles
    // skip the two larger tables
    bool work = false;
    if (MODE != 2 && in) {
        const int64_t cid = m3::cell_id(a.g, c[0], c[1], c[2]);
        work = (a.g.cwork[cid >> 5] >> (cid & 31)) & 1u;
        if (work) {
            it0 = a.g.cstart[cid];
            it1 = a.g.cstart[cid + 1];
        }
    }
    const bool wide = !a.no_faces && it1 - it0 > a.wide_min;
    // the winding number where it is a lookup (outside the grid, a face-free cell); NaN: the walk decides
    double w0 = 0.0;
    bool long_walk = false;
    if (MODE == 0 && a.chi && in) {
        const int64_t cid = m3::cell_id(a.g, c[0], c[1], c[2]);
        const int d = work && a.g.cwalk ? a.g.cwalk[cid] : -1;
        if (d >= 0) {
            w0 = __builtin_nan("");
            long_walk = a.g.wstart[d + 1] - a.g.wstart[d] > a.wide_min;
        } else {
            w0 = a.g.cchi[cid];
        }
    }
    if (MODE == 0) {
        // wide particles go to the list of adj3_wide_kernel, long walks to that of adj3_walk_kernel (one atomic per
        // wavefront and list)
        const int lane = threadIdx.x & 63;
        const uint64_t wm = __ballot(wide);
        if (wm) {
            const int leader = __ffsll((unsigned long long)wm) - 1;
            int first = 0;
            if (lane == leader) first = atomicAdd(a.wide_count, __popcll(wm));
            first = __shfl(first, leader);
            if (wide) a.wide_list[first + __popcll(wm & ((1ull << lane) - 1))] = wi;
        }
        const uint64_t lm = __ballot(long_walk);
        if (lm) {
            const int leader = __ffsll((unsigned long long)lm) - 1;
            int first = 0;
            if (lane == leader) first = atomicAdd(a.walk_count, __popcll(lm));
            first = __shfl(first, leader);
            if (long_walk) a.walk_list[first + __popcll(lm & ((1ull << lane) - 1))] = wi;
        }
    }
    if (MODE != 2 && in && !wide && !a.no_faces) {
        const Near32 nq = near32(px, py, pz, h);
        const int64_t base = FILL ? a.offsets[wi] : 0;
        // ADJ_UNROLL candidates per step: their 32-byte records are loaded together (independent loads in flight),
        // then filtered in list order; the exact test only for boxes within h on every axis
        for (int it = it0; it < it1; it += ADJ_UNROLL) {
            float4 l[ADJ_UNROLL], u[ADJ_UNROLL];
#pragma unroll
            for (int q = 0; q < ADJ_UNROLL; ++q) {
                const int j = min(it + q, it1 - 1);
                l[q] = a.g.crec[2 * (int64_t)j];
                u[q] = a.g.crec[2 * (int64_t)j + 1];
            }
#pragma unroll
            for (int q = 0; q < ADJ_UNROLL; ++q) {
                if (it + q >= it1) break;
                const int k32 = classify32(nq, l[q], u[q]);
                if (k32 == 0) continue;
                const int f = m3::bits_int(l[q].w);
                if (k32 == 2 && !keep_face_fast(a.g.fbox + 6 * (int64_t)f, px, py, pz, h)) continue;
                if (FILL && base + cnt < a.capacity) a.faces[base + cnt] = f;
                ++cnt;
            }
        }
    }
    if (MODE == 1) return;
    if (MODE == 0) {
        if (live) {
            if (!wide && !a.no_faces) a.counts[wi] = cnt;
            if (a.chi) a.chi[wi] = w0;
            if (a.prow) a.prow[wi] = double4{px, py, pz, h};
        }
        return;
    }
    bool need = false;
    if (todo) need = !m3::ray_winding(a.g, px, py, pz, c, &chi);
    const uint64_t mask = __ballot(need);
    if (mask) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int first = 0;
        if (lane == leader) first = atomicAdd(a.exact_count, __popcll(mask));
        first = __shfl(first, leader);
        if (need) a.exact_list[first + __popcll(mask & ((1ull << lane) - 1))] = wi;
    }
    if (todo && !need) a.chi[wi] = chi;          // the listed ones keep NaN until adj3_exact_kernel
}

// face lists from the face tree, one thread per particle: a stack walk of the tree (a node whose FP32 box the
// particle's FP32 filter rejects is skipped whole), the faces of the leaves reached through the same decision as the
// cell lists (FP32 filter, then the element test), so the set is the cell lists' set.  MODE 0: counts; MODE 1: the
// list in ascending face order -- a row of at most BVH_SHORT faces sorted in registers, a longer one written in tree
// order and listed for bvh3_sort_kernel.
constexpr int BVH_SHORT = 32;

template <int MODE>
__global__ __launch_bounds__(256) void bvh3_kernel(Adj3Args a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < a.n;
    const int i = live ? (a.order ? a.order[t] : t) : 0;
    int64_t base = 0, total = 0;
    if (MODE == 1 && live) {
        base = a.offsets[i];
        total = a.offsets[i + 1] - base;
    }
    const bool shortrow = total <= BVH_SHORT;
    const bool fits = base + total <= a.capacity;
    if (MODE == 1) {            // the long rows go to bvh3_sort_kernel (one atomic per wavefront)
        const bool lng = live && !shortrow && fits;
        const uint64_t lm = __ballot(lng);
        if (lm) {
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((unsigned long long)lm) - 1;
            int first = 0;
            if (lane == leader) first = atomicAdd(a.long_count, __popcll(lm));
            first = __shfl(first, leader);
            if (lng) a.long_list[first + __popcll(lm & ((1ull << lane) - 1))] = i;
        }
    }
    if (!live) return;
    const double px = a.pos[3 * (int64_t)i], py = a.pos[3 * (int64_t)i + 1], pz = a.pos[3 * (int64_t)i + 2];
    const double h = a.sup[i];
    if (MODE == 0 && !(h <= a.hlim)) atomicOr(a.bad, 1);
    const Near32 nq = near32(px, py, pz, h);
    int32_t buf[BVH_SHORT];
    int64_t cnt = 0;
    int stack[BVH_STACK];
    int sp = 0, node = 0;
    for (;;) {
        const float4 lo = a.bnode[2 * node], hi = a.bnode[2 * node + 1];
        bool go = classify32(nq, lo, hi) != 0;
        if (go) {
            const int l = m3::bits_int(lo.w), r = m3::bits_int(hi.w);
            if (l >= 0) {
                stack[sp++] = r;
                node = l;
                continue;
            }
            const int first = -l - 1;
            for (int it = first; it < first + r; ++it) {
                const float4 il = a.bitem[2 * (int64_t)it], iu = a.bitem[2 * (int64_t)it + 1];
                const int k32 = classify32(nq, il, iu);
                if (k32 == 0) continue;
                const int f = m3::bits_int(il.w);
                if (k32 == 2 && !keep_face_fast(a.g.fbox + 6 * (int64_t)f, px, py, pz, h)) continue;
                if (MODE == 1 && fits) {
                    if (shortrow) {
                        if (cnt < BVH_SHORT) buf[cnt] = f;
                    } else {
                        a.faces[base + cnt] = f;
                    }
                }
                ++cnt;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    if (MODE == 0) {
        a.counts[i] = cnt;
        return;
    }
    if (shortrow && fits) {
        const int c = (int)min(cnt, (int64_t)BVH_SHORT);
        for (int p = 1; p < c; ++p) {                  // insertion sort: ascending face ids
            const int32_t v = buf[p];
            int q = p - 1;
            while (q >= 0 && buf[q] > v) {
                buf[q + 1] = buf[q];
                --q;
            }
            buf[q + 1] = v;
        }
        for (int p = 0; p < c; ++p) a.faces[base + p] = buf[p];
    }
}

// the long rows listed by bvh3_kernel<1>, one workgroup each: the row's face ids marked in an LDS bitmap over all
// face ids, then written back in ascending order (the bits of each thread's words, placed after a scan of the
// threads' counts).  Faces of a mesh: at most BVH_BITS.
constexpr int BVH_WORDS = 8192;                        // 32 KiB of LDS
constexpr int BVH_BITS = BVH_WORDS * 32;

__global__ __launch_bounds__(256) void bvh3_sort_kernel(Adj3Args a) {
    using BlockScan = hipcub::BlockScan<int, 256>;
    __shared__ typename BlockScan::TempStorage scan_tmp;
    __shared__ uint32_t bits[BVH_WORDS];
    const int nlong = *a.long_count;
    const int nw = (a.g.nf + 31) / 32;
    constexpr int PER = BVH_WORDS / 256;
    for (int q = blockIdx.x; q < nlong; q += gridDim.x) {
        const int i = a.long_list[q];
        const int64_t base = a.offsets[i];
        const int total = (int)(a.offsets[i + 1] - base);
        for (int w = threadIdx.x; w < nw; w += 256) bits[w] = 0u;
        __syncthreads();
        for (int j = threadIdx.x; j < total; j += 256) {
            const int f = a.faces[base + j];
            atomicOr(&bits[f >> 5], 1u << (f & 31));
        }
        __syncthreads();
        const int w0 = threadIdx.x * PER;
        int mine = 0;
        for (int w = w0; w < w0 + PER && w < nw; ++w) mine += __popc(bits[w]);
        int at = 0;
        BlockScan(scan_tmp).ExclusiveSum(mine, at);
        for (int w = w0; w < w0 + PER && w < nw; ++w) {
            uint32_t v = bits[w];
            while (v) {
                const int b = __ffs((int)v) - 1;
                a.faces[base + at++] = w * 32 + b;
                v &= v - 1u;
            }
        }
        __syncthreads();                            // bits and scan_tmp are reused by the next row
    }
}

// the particles listed by adj3_kernel<0> for their long axis walk, one wavefront each: lane k takes walk items k,
// k + 64, ...; any item within its rounding bound sends the particle to the exact per-face sum (as the one-lane walk
// does).  Writes the winding number, so adj3_kernel<2> passes these particles by.
__global__ __launch_bounds__(256) void adj3_walk_kernel(Adj3Args a) {
    const int lane = threadIdx.x & 63;
    const int64_t wave_id = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int nwalk = *a.walk_count;
    for (int64_t q = wave_id; q < nwalk; q += nwaves) {
        const int wi = a.walk_list[q], i = adj3_row(a, wi);
        const double4 pr = a.prow ? a.prow[wi]
                                  : double4{a.pos[3 * (int64_t)i], a.pos[3 * (int64_t)i + 1], a.pos[3 * (int64_t)i + 2], a.sup[i]};
        const double p[3] = {pr.x, pr.y, pr.z};
        const double h = pr.w;
        int c[3] = {0, 0, 0};
        (void)m3::cell_of(a.g, p[0], p[1], p[2], c);       // inside the grid: it was listed from its cell
        const m3::Walk w = m3::walk_of(a.g, c, a.g.cwalk[m3::cell_id(a.g, c[0], c[1], c[2])]);
        const double Pa = p[w.ax], Pu = p[w.au], Pv = p[w.av];
        int net = 0;
        bool unsure = false;
        for (int it = w.it0; it < w.it1; it += 64) {
            const int j = it + lane;
            int r = 0;
            if (j < w.it1)
                r = m3::walk_item(a.g.frec, p[0], p[1], p[2], Pa, Pu, Pv, w.au, w.av, w.s, a.g.wrec[2 * (int64_t)j],
                                  a.g.wrec[2 * (int64_t)j + 1]);
            unsure = unsure || r == 2;
            net += (r == 2) ? 0 : r;
        }
        if (__ballot(unsure)) {
            if (lane == 0) a.exact_list[atomicAdd(a.exact_count, 1)] = wi;    // NaN until adj3_exact_kernel
        } else {
            for (int o = 32; o > 0; o >>= 1) net += __shfl_xor(net, o);
            if (lane == 0) a.chi[wi] = w.acc + (double)net;
        }
    }
}

// the reference's per-face sum for the particles whose walk hit a rounding bound (listed by adj3_kernel<2> and
// adj3_walk_kernel), one wavefront each
__global__ __launch_bounds__(256) void adj3_exact_kernel(Adj3Args a) {
    const int lane = threadIdx.x & 63;
    const int64_t wave_id = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int nexact = *a.exact_count;
    for (int64_t q = wave_id; q < nexact; q += nwaves) {
        const int wi = a.exact_list[q], i = adj3_row(a, wi);
        const double4 pr = a.prow ? a.prow[wi]
                                  : double4{a.pos[3 * (int64_t)i], a.pos[3 * (int64_t)i + 1], a.pos[3 * (int64_t)i + 2], a.sup[i]};
        const double s = wave_exact_winding(a.g, pr.x, pr.y, pr.z, 1.0 / pr.w);
        if (lane == 0) a.chi[wi] = s;
    }
}

// the particles listed by adj3_kernel<0> as wide, one wavefront each, a fixed grid of wavefronts taking them in turn:
// FILL false writes the pair counts, true the face lists
template <bool FILL>
__global__ __launch_bounds__(256) void adj3_wide_kernel(Adj3Args a) {
    const int lane = threadIdx.x & 63;
    const int64_t wave_id = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int nwide = *a.wide_count;
    const uint64_t below = (1ull << lane) - 1;
    for (int64_t q = wave_id; q < nwide; q += nwaves) {
        const int wi = a.wide_list[q], i = adj3_row(a, wi);
        const double4 pr = a.prow ? a.prow[wi]
                                  : double4{a.pos[3 * (int64_t)i], a.pos[3 * (int64_t)i + 1], a.pos[3 * (int64_t)i + 2], a.sup[i]};
        const double px = pr.x, py = pr.y, pz = pr.z;
        const double h = pr.w;
        int c[3] = {0, 0, 0};
        (void)m3::cell_of(a.g, px, py, pz, c);        // inside the grid: it was listed from its cell
        const int64_t cid = m3::cell_id(a.g, c[0], c[1], c[2]);
        const int it0 = a.g.cstart[cid], it1 = a.g.cstart[cid + 1];
        const Near32 nq = near32(px, py, pz, h);
        const int64_t base = FILL ? a.offsets[wi] : 0;
        int64_t cnt = 0;
        for (int it = it0; it < it1; it += 64) {
            const int j = it + lane;
            bool keep = false;
            int f = 0;
            if (j < it1) {
                const float4 l = a.g.crec[2 * (int64_t)j], u = a.g.crec[2 * (int64_t)j + 1];
                const int k32 = classify32(nq, l, u);
                f = m3::bits_int(l.w);
                keep = k32 == 1 || (k32 == 2 && keep_face_fast(a.g.fbox + 6 * (int64_t)f, px, py, pz, h));
            }
            const uint64_t m = __ballot(keep);
            if (FILL && keep) {
                const int64_t p = base + cnt + __popcll(m & below);
                if (p < a.capacity) a.faces[p] = f;
            }
            cnt += __popcll(m);
        }
        if (!FILL && lane == 0) a.counts[wi] = cnt;
    }
}

__global__ void offsets_head_kernel(int64_t* offsets) { offsets[0] = 0; }

// counting sort of the particles that have work by grid cell: histogram, exclusive scan, scatter (the order inside
// a cell depends on the atomics' timing; it only decides which thread serves a particle, never a result).  A
// particle has work when its cell lists candidate faces or has faces of its own (the axis walk); every other
// particle (outside the grid, or in a face-free cell with no face within support_max) keeps its input order after
// the sorted ones, placed by a scan of their flags, so their loads and stores stay coalesced.
__global__ __launch_bounds__(256) void cell_count3_kernel(m3::Grid3 g, const double* pos, int n, int32_t* key, int32_t* outflag,
                                                          int32_t* hist) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int c[3];
    const bool in = m3::cell_of(g, pos[3 * (int64_t)r], pos[3 * (int64_t)r + 1], pos[3 * (int64_t)r + 2], c);
    int k = -1;
    if (in) {
        const int64_t cid = m3::cell_id(g, c[0], c[1], c[2]);
        if ((g.cwork[cid >> 5] >> (cid & 31)) & 1u) k = (int)cid;
    }
    key[r] = k;
    outflag[r] = k >= 0 ? 0 : 1;
    if (k >= 0) atomicAdd(hist + k, 1);
}

__global__ __launch_bounds__(256) void cell_scatter3_kernel(const int32_t* key, const int32_t* outflag, const int32_t* outpos, int n,
                                                            int32_t* cursor, int32_t* order) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int k = key[r];
    if (k >= 0) {
        order[atomicAdd(cursor + k, 1)] = r;
    } else {
        const int n_out = outpos[n - 1] + outflag[n - 1];
        order[(n - n_out) + outpos[r]] = r;
    }
}

// the key pass of the two-pass sort in one launch: every particle's key -- its cell when the cell's bit is set in the
// work bitmap (candidate faces or a walk), else -1 (placed after the others) -- from the positions of the chunk, all
// loaded before any key is formed, and the chunk's LDS histogram of coarse bins as cs_hist_keys_kernel
__global__ __launch_bounds__(CS_T) void cs_hist3_kernel(m3::Grid3 g, const double* pos, int n, int shift, int nc, int nblk,
                                                        int32_t* key, int32_t* ghist) {
    __shared__ int hs[CS_NC];
    for (int i = threadIdx.x; i < nc; i += CS_T) hs[i] = 0;
    const int64_t c0 = (int64_t)blockIdx.x * CS_CH;
    double px[CS_RPT], py[CS_RPT], pz[CS_RPT];
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        const bool live = r < n;
        px[j] = live ? pos[3 * r] : 0.0;
        py[j] = live ? pos[3 * r + 1] : 0.0;
        pz[j] = live ? pos[3 * r + 2] : 0.0;
    }
    int k[CS_RPT];
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        int c[3];
        k[j] = -1;
        if (r < n && m3::cell_of(g, px[j], py[j], pz[j], c)) {
            const int64_t cid = m3::cell_id(g, c[0], c[1], c[2]);
            if ((g.cwork[cid >> 5] >> (cid & 31)) & 1u) k[j] = (int)cid;
        }
        if (r < n) key[r] = k[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const bool live = c0 + j * CS_T + threadIdx.x < n;
        const uint64_t out = __ballot(live && k[j] < 0);
        if (live && k[j] >= 0) atomicAdd(&hs[k[j] >> shift], 1);
        else if (live && lane_rank(out) == 0) atomicAdd(&hs[nc - 1], __popcll(out));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += CS_T) ghist[(int64_t)i * nblk + blockIdx.x] = hs[i];
}

__global__ void zero3_kernel(int32_t* p, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

struct Sort3Layout {
    size_t key, flag, opos, hist, cur, tmp, tmp_bytes, total;
};

inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

Sort3Layout sort3_layout(int n, int64_t nbins) {
    Sort3Layout L{};
    size_t tmp = 0, tmp2 = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const int32_t*)nullptr, (int32_t*)nullptr, (int)nbins);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, (const int32_t*)nullptr, (int32_t*)nullptr, n);
    tmp = std::max(tmp, tmp2);
    L.key = 0;
    L.flag = al256(L.key + sizeof(int32_t) * (size_t)n);
    L.opos = al256(L.flag + sizeof(int32_t) * (size_t)n);
    L.hist = al256(L.opos + sizeof(int32_t) * (size_t)n);
    L.cur = al256(L.hist + sizeof(int32_t) * (size_t)nbins);
    L.tmp = al256(L.cur + sizeof(int32_t) * (size_t)nbins);
    L.tmp_bytes = tmp;
    L.total = al256(L.tmp + tmp);
    return L;
}

// ------------------------------------------------------------------------------------------------ moment tiles
constexpr int TILE3 = 256;

struct Mom3Args {
    const int64_t* offsets;
    const int32_t* faces;
    int64_t capacity;
    const double* chi;
    const double* pos;
    const double* sup;
    int n;
    const double* frec;
    int nf;
    int flags;
    double* out;
    int32_t* heavy_count;       // scratch: the number of heavy particles, then their row indices
    int32_t* heavy_list;
    int rows_per_tile;          // particles per tile22_kernel workgroup, at most TILE3
    const double4* prow;        // non-null: the sorted-position layout (offsets, faces, chi and (x, y, z, h) by position;
    const int32_t* order;       // the particle at position i is order[i], whose row of out it fills)
};

// A particle with more than HEAVY3 pairs is served by one workgroup (heavy22_kernel): pair j of it goes to thread
// j % 256, each thread sums its pairs in order, a butterfly over each wavefront's 64 lanes adds the thread sums (every
// lane ends with the same value: each level adds the same two operands) and the four wavefront sums are added in a
// fixed order.  The other particles are served in tiles (tile22_kernel): a
// workgroup of TILE3 threads owns rows_per_tile <= TILE3 consecutive particles (fewer when the launch is small, so
// that it still has thousands of workgroups), their pairs are packed densely over the threads (pair j -> its particle
// by a binary search of the tile's offsets in LDS); every thread evaluates one pair, the 23 face accumulators go
// through LDS and each particle's thread adds its own pairs in face order.  Either way a particle's sums depend only on its own
// pairs, never on the packing or on the other particles of its batch.
constexpr int HEAVY3 = 64;

// MODE 0: the 22 stored channels (23 face accumulators, winding term); MODE 1: the six Hessian entries (six
// accumulators, no winding term: the whole ball's Hessian is zero)
template <int MODE> struct Mode3;
template <> struct Mode3<0> {
    static constexpr int NA = m3::NACC, NO = m3::NOUT;
    template <int KID, int FORM>
    __device__ static void pair(const double* fr, double px, double py, double pz, double ih, double* acc) {
        m3::face_terms<KID, FORM>(fr, px, py, pz, ih, acc);
    }
    template <int KID>
    __device__ static void finalize(const double* acc, double chi, double h, bool physical, double* out) {
        m3::finalize22<KID>(acc, chi, h, physical, out);
    }
};
template <> struct Mode3<1> {
    static constexpr int NA = m3::NHESS, NO = m3::NHESS;
    template <int KID, int FORM>
    __device__ static void pair(const double* fr, double px, double py, double pz, double ih, double* acc) {
        m3::face_hessian<KID, FORM>(fr, px, py, pz, ih, acc);
    }
    template <int KID>
    __device__ static void finalize(const double* acc, double, double h, bool physical, double* out) {
        m3::finalize_hessian(acc, h, physical, out);
    }
};

template <int KID, int FORM, int MODE>
__global__ __launch_bounds__(TILE3) void tile22_kernel(Mom3Args a) {
    using M = Mode3<MODE>;
    using BlockScan = hipcub::BlockScan<int, TILE3>;
    __shared__ typename BlockScan::TempStorage scan_tmp;
    __shared__ int s_off[TILE3 + 1];
    __shared__ int s_first[TILE3];
    __shared__ int s_ord[TILE3];          // the sorted layout: the row each position of the tile fills
    __shared__ double s_p[4][TILE3];
    __shared__ double s_res[M::NA][TILE3 + 1];     // one spare column: the store loop's lanes read one row's values
                                                   // (consecutive k), TILE3 + 1 doubles apart, in distinct banks
    const int tid = threadIdx.x;
    const int64_t r0 = blockIdx.x * (int64_t)a.rows_per_tile;
    const int64_t row = r0 + tid;
    const bool live = tid < a.rows_per_tile && row < a.n;
    int cnt = 0;
    double h = 1.0;
    if (live) {
        cnt = (int)(a.offsets[row + 1] - a.offsets[row]);
        if (a.prow) {
            const double4 pr = a.prow[row];
            s_p[0][tid] = pr.x;
            s_p[1][tid] = pr.y;
            s_p[2][tid] = pr.z;
            h = pr.w;
            s_ord[tid] = a.order[row];
        } else {
            s_p[0][tid] = a.pos[3 * row];
            s_p[1][tid] = a.pos[3 * row + 1];
            s_p[2][tid] = a.pos[3 * row + 2];
            h = a.sup[row];
        }
        // a heavy particle (served by heavy22_kernel) is marked by a negative inverse support in LDS
        s_p[3][tid] = cnt > HEAVY3 ? -1.0 : 1.0 / h;
    }
    const bool heavy = cnt > HEAVY3;
    if (heavy) a.heavy_list[atomicAdd(a.heavy_count, 1)] = (int32_t)row;
    // heavy particles' pairs are left out of the packing: then pair j of the tile sits at slot tile_first +
    // s_first[r] + (j - s_off[r]) for its row r; a tile without heavy particles keeps slot tile_first + j
    const int64_t tile_first = r0 < a.n ? a.offsets[r0] : 0;
    if (live) s_first[tid] = (int)(a.offsets[row] - tile_first);
    int off, total;
    BlockScan(scan_tmp).ExclusiveSum(heavy ? 0 : cnt, off, total);
    s_off[tid] = off;
    if (tid == 0) s_off[TILE3] = total;
    const bool mapped = __syncthreads_or(heavy);
    double acc[M::NA];
#pragma unroll
    for (int k = 0; k < M::NA; ++k) acc[k] = 0.0;
    for (int base = 0; base < total; base += TILE3) {
        const int j = base + tid;
        double c[M::NA];
#pragma unroll
        for (int k = 0; k < M::NA; ++k) c[k] = 0.0;
        if (j < total) {
            int lo = 0, hi = TILE3;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= j) lo = mid;
                else hi = mid;
            }
            const int64_t slot = tile_first + (mapped ? s_first[lo] + (j - s_off[lo]) : j);
            const int f = slot < a.capacity ? a.faces[slot] : -1;
            const double ihl = s_p[3][lo];
            if (f >= 0 && f < a.nf && ihl > 0.0)
                M::template pair<KID, FORM>(a.frec + m3::FREC * (int64_t)f, s_p[0][lo], s_p[1][lo], s_p[2][lo], ihl, c);
        }
#pragma unroll
        for (int k = 0; k < M::NA; ++k) s_res[k][tid] = c[k];
        __syncthreads();
        const int b0 = max(off, base), b1 = heavy ? b0 : min(off + cnt, base + TILE3);   // off + 0 for heavy
        for (int jj = b0; jj < b1; ++jj) {
#pragma unroll
            for (int k = 0; k < M::NA; ++k) acc[k] += s_res[k][jj - base];
        }
        __syncthreads();
    }
    // finalise into LDS, then store the tile's rows as one contiguous block (coalesced across the lanes)
    if (live) {
        double o[M::NO];
        const double chi = (MODE != 0 || (a.flags & WSB_OUT_NO_INDICATOR)) ? 0.0 : a.chi[row];
        M::template finalize<KID>(acc, chi, h, (a.flags & WSB_OUT_PHYSICAL) != 0, o);
#pragma unroll
        for (int k = 0; k < M::NO; ++k) s_res[k][tid] = o[k];
    }
    __syncthreads();
    const int nrows = (int)min((int64_t)a.rows_per_tile, (int64_t)a.n - r0);
    const bool accumulate = (a.flags & WSB_OUT_ACCUMULATE) != 0;
    for (int idx = tid; idx < nrows * M::NO; idx += TILE3) {
        const int rr = idx / M::NO, k = idx - rr * M::NO;
        if (s_p[3][rr] < 0.0) continue;                // stored by heavy22_kernel
        const double v = s_res[k][rr];
        // the sorted layout's rows lie apart, each stored by consecutive lanes (its NO values contiguous)
        double* q = a.out + (a.order ? (int64_t)s_ord[rr] * M::NO + k : r0 * M::NO + idx);
        if (accumulate) *q += v;
        else *q = v;
    }
}

// the particles with more than HEAVY3 pairs (listed by tile22_kernel), one workgroup of four wavefronts each (a
// fixed grid of workgroups taking them in turn): pair j of a particle goes to thread j % 256, each thread sums its
// pairs in order, a butterfly adds the 64 sums of each wavefront, and the four wavefront sums are added in wavefront
// order -- the same order whatever else is in the batch
template <int KID, int FORM, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void heavy22_kernel(Mom3Args a) {
    using M = Mode3<MODE>;
    __shared__ double s_acc[4][M::NA];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const bool accumulate = (a.flags & WSB_OUT_ACCUMULATE) != 0;
    const int nheavy = *a.heavy_count;
    for (int64_t q = blockIdx.x; q < nheavy; q += gridDim.x) {
        const int64_t r = a.heavy_list[q];
        const int64_t first = a.offsets[r];
        const int rc = (int)(a.offsets[r + 1] - first);
        double px, py, pz, h;
        if (a.prow) {
            const double4 pr = a.prow[r];
            px = pr.x;
            py = pr.y;
            pz = pr.z;
            h = pr.w;
        } else {
            px = a.pos[3 * r];
            py = a.pos[3 * r + 1];
            pz = a.pos[3 * r + 2];
            h = a.sup[r];
        }
        const double ih = 1.0 / h;
        double acc[M::NA];
#pragma unroll
        for (int k = 0; k < M::NA; ++k) acc[k] = 0.0;
        for (int j = tid; j < rc; j += 256) {
            const int64_t slot = first + j;
            const int f = slot < a.capacity ? a.faces[slot] : -1;
            if (f >= 0 && f < a.nf) M::template pair<KID, FORM>(a.frec + m3::FREC * (int64_t)f, px, py, pz, ih, acc);
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
            for (int k = 0; k < M::NA; ++k) acc[k] += __shfl_xor(acc[k], o);
        }
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < M::NA; ++k) s_acc[wv][k] = acc[k];
        }
        __syncthreads();
        if (wv == 0) {
            double tot[M::NA];
#pragma unroll
            for (int k = 0; k < M::NA; ++k) tot[k] = ((s_acc[0][k] + s_acc[1][k]) + s_acc[2][k]) + s_acc[3][k];
            double out[M::NO];
            const double chi = (MODE != 0 || (a.flags & WSB_OUT_NO_INDICATOR)) ? 0.0 : a.chi[r];
            M::template finalize<KID>(tot, chi, h, (a.flags & WSB_OUT_PHYSICAL) != 0, out);
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < M::NO; ++k) v = (k == lane) ? out[k] : v;
            if (lane < M::NO) {
                double* dst = a.out + (a.order ? (int64_t)a.order[r] : r) * M::NO + lane;
                if (accumulate) *dst += v;
                else *dst = v;
            }
        }
        __syncthreads();            // s_acc is reused by the next particle
    }
}

__global__ void zero_count3_kernel(int32_t* c) { *c = 0; }
__global__ void zero3_counts_kernel(int32_t* c) {
    c[0] = 0;
    c[1] = 0;
    c[2] = 0;
    c[3] = 0;
    c[4] = 0;
}

// a support above the mesh's support_max: the pair total offsets[n] reads -1 (the lists are not valid)
__global__ void adj3_mark_kernel(const int32_t* bad, int64_t* total) {
    if (*bad) *total = -1;
}

template <int KID, int FORM, int MODE>
int launch22(const Mom3Args& a, hipStream_t s) {
    hipLaunchKernelGGL(zero_count3_kernel, dim3(1), dim3(1), 0, s, a.heavy_count);
    // rows per tile: TILE3, halved (down to 16) while the launch has fewer than 2,048 workgroups
    Mom3Args t = a;
    // sorted layout: 128 particles per workgroup (grouped by cell, their pairs are dense; 64 or 256 measured slower on
    // the sphere cavity)
    t.rows_per_tile = a.prow ? TILE3 / 2 : TILE3;
    while (t.rows_per_tile > 16 && ((int64_t)a.n + t.rows_per_tile - 1) / t.rows_per_tile < 2048) t.rows_per_tile >>= 1;
    hipLaunchKernelGGL((tile22_kernel<KID, FORM, MODE>), dim3((unsigned)(((int64_t)a.n + t.rows_per_tile - 1) / t.rows_per_tile)),
                       dim3(TILE3), 0, s, t);
    WSB_HIP(hipGetLastError());
    // up to 2,048 workgroups (8 per CU), one per listed particle at a time
    const int blocks = (int)std::min<int64_t>((int64_t)a.n, 2048);
    hipLaunchKernelGGL((heavy22_kernel<KID, FORM, MODE>), dim3(blocks), dim3(256), 0, s, a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

template <int FORM, int MODE = 0>
int dispatch22(int kid, const Mom3Args& a, hipStream_t s) {
    switch (kid) {
        case WSB_CUBIC: return launch22<WSB_CUBIC, FORM, MODE>(a, s);
        case WSB_W2: return launch22<WSB_W2, FORM, MODE>(a, s);
        case WSB_W4: return launch22<WSB_W4, FORM, MODE>(a, s);
        case WSB_W6: return launch22<WSB_W6, FORM, MODE>(a, s);
    }
    return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kid);
}

template <int KID, int FORM>
void host_pairs(const double* frec, const int32_t* pf, const double* pq, const double* ph, int P, double* out) {
    for (int j = 0; j < P; ++j) {
        double c[m3::NACC];
        for (int k = 0; k < m3::NACC; ++k) c[k] = 0.0;
        m3::face_terms<KID, FORM>(frec + m3::FREC * (int64_t)pf[j], pq[3 * j], pq[3 * j + 1], pq[3 * j + 2], 1.0 / ph[j], c);
        for (int k = 0; k < m3::NACC; ++k) out[(int64_t)j * m3::NACC + k] = c[k];
    }
}

template <int KID, int FORM>
void host_hessian_pairs(const double* frec, const int32_t* pf, const double* pq, const double* ph, int P, double* out) {
    for (int j = 0; j < P; ++j) {
        double c[m3::NHESS];
        for (int k = 0; k < m3::NHESS; ++k) c[k] = 0.0;
        m3::face_hessian<KID, FORM>(frec + m3::FREC * (int64_t)pf[j], pq[3 * j], pq[3 * j + 1], pq[3 * j + 2], 1.0 / ph[j], c);
        for (int k = 0; k < m3::NHESS; ++k) out[(int64_t)j * m3::NHESS + k] = c[k];
    }
}

template <int KID>
void host_finalize(const double* acc, const double* chi, const double* h, int n, int flags, double* out) {
    for (int i = 0; i < n; ++i) m3::finalize22<KID>(acc + (int64_t)i * m3::NACC, chi[i], h[i], (flags & WSB_OUT_PHYSICAL) != 0, out + (int64_t)i * m3::NOUT);
}

}  // namespace

// ================================================================================================= C ABI
extern "C" int wsb_mesh3d_create(const double* verts, int nv, const int32_t* faces, int nf, double support_max, int flags,
                                 wsb_mesh3d** out, int64_t* info) {
    WSB_REQUIRE(verts && faces && out, WSB_EINVAL, "null argument");
    WSB_REQUIRE(nv >= 3 && nf >= 1, WSB_EINVAL, "closed surface needs faces (nv=%d nf=%d)", nv, nf);
    WSB_REQUIRE(support_max > 0 && std::isfinite(support_max), WSB_EINVAL, "support_max must be positive and finite");
    for (int64_t i = 0; i < 3 * (int64_t)nv; ++i) WSB_REQUIRE(std::isfinite(verts[i]), WSB_EINVAL, "non-finite vertex");
    for (int f = 0; f < nf; ++f) {
        const int32_t* ix = faces + 3 * (int64_t)f;
        for (int k = 0; k < 3; ++k) WSB_REQUIRE(ix[k] >= 0 && ix[k] < nv, WSB_EINVAL, "invalid triangle indices (face %d)", f);
        WSB_REQUIRE(ix[0] != ix[1] && ix[1] != ix[2] && ix[0] != ix[2], WSB_EINVAL, "invalid triangle indices (face %d)", f);
    }
    const bool soup = (flags & WSB_MESH_SOUP) != 0;
    if (!soup) {
        const int st = validate_closed(verts, nv, faces, nf);
        if (st) return st;
    }
    auto* m = new wsb_mesh3d();
    m->host_only = (flags & WSB_MESH_HOST_ONLY) != 0;
    m->soup = soup;
    m->nv = nv;
    m->nf = nf;
    m->support_max = support_max;
    int st = build_host(m, verts, nv, faces, nf, support_max);
    if (st) {
        delete m;
        return st;
    }
    if (soup) m->hg.cwalk = nullptr;
    {
        const int64_t ncells = (int64_t)m->hg.n[0] * m->hg.n[1] * m->hg.n[2];
        m->cwork.assign((size_t)((ncells + 31) / 32), 0u);
        for (int64_t c = 0; c < ncells; ++c)
            if (m->cstart[c + 1] > m->cstart[c] || (!soup && m->cwalk[c] >= 0)) m->cwork[c >> 5] |= 1u << (c & 31);
        m->hg.cwork = m->cwork.data();
    }
    m->dg = m->hg;
    if (!m->host_only) {
        if (hipGetDevice(&m->device) != hipSuccess) {
            delete m;
            return wsb::set_error(WSB_EDEVICE, "no HIP device");
        }
        if (!st) st = upload(&m->d_frec, m->frec, &m->bytes);
        if (!st) st = upload(&m->d_fbox, m->fbox, &m->bytes);
        if (!st) st = upload(&m->d_cstart, m->cstart, &m->bytes);
        if (!st) st = upload(&m->d_citems, m->citems, &m->bytes);
        if (!st) st = upload(&m->d_crec, m->crec, &m->bytes);
        if (!st) st = upload(&m->d_cchi, m->cchi, &m->bytes);
        if (!st) st = upload(&m->d_cwalk, m->cwalk, &m->bytes);
        if (!st) st = upload(&m->d_cwork, m->cwork, &m->bytes);
        if (!st) st = upload(&m->d_wdir, m->wdir, &m->bytes);
        if (!st) st = upload(&m->d_wstart, m->wstart, &m->bytes);
        if (!st) st = upload(&m->d_witems, m->witems, &m->bytes);
        if (!st) st = upload(&m->d_wrec, m->wrec, &m->bytes);
        if (!st) st = build_bvh(m);
        if (!st) st = upload(&m->d_bnode, m->bnode, &m->bytes);
        if (!st) st = upload(&m->d_bitem, m->bitem, &m->bytes);
        if (st) {
            wsb_mesh3d_destroy(m);
            return st;
        }
        m->dg.frec = m->d_frec;
        m->dg.fbox = m->d_fbox;
        m->dg.cstart = m->d_cstart;
        m->dg.citems = m->d_citems;
        m->dg.crec = m->d_crec;
        m->dg.cchi = m->d_cchi;
        m->dg.cwalk = soup ? nullptr : m->d_cwalk;
        m->dg.cwork = m->d_cwork;
        m->dg.wdir = m->d_wdir;
        m->dg.wstart = m->d_wstart;
        m->dg.witems = m->d_witems;
        m->dg.wrec = m->d_wrec;
    }
    *out = m;
    if (info) wsb_mesh3d_info(m, info);
    return WSB_OK;
}

extern "C" int wsb_mesh3d_destroy(wsb_mesh3d* m) {
    if (!m) return WSB_OK;
    for (void* p : {(void*)m->d_frec, (void*)m->d_fbox, (void*)m->d_cstart, (void*)m->d_citems, (void*)m->d_crec, (void*)m->d_cchi,
                    (void*)m->d_cwalk, (void*)m->d_cwork,
                    (void*)m->d_wdir, (void*)m->d_wstart, (void*)m->d_witems, (void*)m->d_wrec, (void*)m->d_bnode,
                    (void*)m->d_bitem})
        if (p) (void)hipFree(p);
    delete m;
    return WSB_OK;
}

extern "C" int wsb_mesh3d_info(const wsb_mesh3d* m, int64_t* info) {
    WSB_REQUIRE(m && info, WSB_EINVAL, "null argument");
    info[0] = m->K;
    info[1] = m->hg.n[0];
    info[2] = m->hg.n[1];
    info[3] = m->hg.n[2];
    info[4] = m->nf;
    info[5] = m->nitems;
    info[6] = m->bytes;
    info[7] = m->ncomp;
    info[8] = m->ndirty;
    int64_t wmax = 0;
    for (size_t d = 0; d + 1 < m->wstart.size(); ++d) wmax = std::max<int64_t>(wmax, m->wstart[d + 1] - m->wstart[d]);
    info[9] = wmax;
    return WSB_OK;
}

extern "C" int wsb_mesh3d_grid(const wsb_mesh3d* m, double* lo_cell) {
    WSB_REQUIRE(m && lo_cell, WSB_EINVAL, "null argument");
    lo_cell[0] = m->hg.lo[0];
    lo_cell[1] = m->hg.lo[1];
    lo_cell[2] = m->hg.lo[2];
    lo_cell[3] = m->hg.cell;
    return WSB_OK;
}

extern "C" size_t wsb_sort3d_scratch_bytes(const wsb_mesh3d* m, int n) {
    if (!m || n <= 0) return 0;
    const int64_t nbins = (int64_t)m->hg.n[0] * m->hg.n[1] * m->hg.n[2];
    return two_pass_sort(nbins) ? cs_layout(n, nbins).total : sort3_layout(n, nbins).total;
}

extern "C" int wsb_sort3d(const wsb_mesh3d* m, const double* pos, int n, int32_t* order, void* scratch, size_t scratch_bytes,
                          void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(!m->host_only, WSB_EINVAL, "host-only mesh has no device tables");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(pos && order, WSB_EINVAL, "null buffer");
    const int64_t nbins = (int64_t)m->hg.n[0] * m->hg.n[1] * m->hg.n[2];
    WSB_REQUIRE(nbins < (int64_t)1 << 31, WSB_ENOTSUP, "grid too large for the cell sort");
    char* base = static_cast<char*>(scratch);
    if (two_pass_sort(nbins)) {
        // the cell keys, then the two-pass counting sort without global atomics (cellsort.h)
        const CsLayout C = cs_layout(n, nbins);
        WSB_REQUIRE(scratch && scratch_bytes >= C.total, WSB_ENOMEM, "sort needs %zu scratch bytes (got %zu)", C.total,
                    scratch_bytes);
        hipStream_t s = wsb::as_stream(stream);
        int32_t* key = reinterpret_cast<int32_t*>(base + C.key);
        const CoarseGrid cg = coarse_grid(nbins);
        const int nblk = (n + CS_CH - 1) / CS_CH;
        hipLaunchKernelGGL(cs_hist3_kernel, dim3(nblk), dim3(CS_T), 0, s, m->dg, pos, n, cg.shift, cg.ncb + 1, nblk, key,
                           reinterpret_cast<int32_t*>(base + C.ghist));
        WSB_HIP(hipGetLastError());
        return cs_place(C, base, n, nbins, order, s);
    }
    const Sort3Layout L = sort3_layout(n, nbins);
    WSB_REQUIRE(scratch && scratch_bytes >= L.total, WSB_ENOMEM, "sort needs %zu scratch bytes (got %zu)", L.total, scratch_bytes);
    int32_t* key = reinterpret_cast<int32_t*>(base + L.key);
    int32_t* flag = reinterpret_cast<int32_t*>(base + L.flag);
    int32_t* opos = reinterpret_cast<int32_t*>(base + L.opos);
    int32_t* hist = reinterpret_cast<int32_t*>(base + L.hist);
    int32_t* cur = reinterpret_cast<int32_t*>(base + L.cur);
    hipStream_t s = wsb::as_stream(stream);
    hipLaunchKernelGGL(zero3_kernel, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, s, hist, nbins);
    hipLaunchKernelGGL(cell_count3_kernel, dim3((n + 255) / 256), dim3(256), 0, s, m->dg, pos, n, key, flag, hist);
    WSB_HIP(hipGetLastError());
    size_t tmp = L.tmp_bytes;
    WSB_HIP(hipcub::DeviceScan::ExclusiveSum(base + L.tmp, tmp, hist, cur, (int)nbins, s));
    tmp = L.tmp_bytes;
    WSB_HIP(hipcub::DeviceScan::ExclusiveSum(base + L.tmp, tmp, flag, opos, n, s));
    hipLaunchKernelGGL(cell_scatter3_kernel, dim3((n + 255) / 256), dim3(256), 0, s, key, flag, opos, n, cur, order);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

// inclusive scan of the pair counts into offsets[1..n] (int64): per chunk of SCAN_CH counts its sum, one workgroup
// scanning the chunk sums (exclusive, in place), then per chunk the scan through LDS -- loaded and stored coalesced, a
// thread scanning its run of SCAN_RPT counts, the element e at e + e / SCAN_RPT so that the runs fall in distinct banks
constexpr int SCAN_T = 256, SCAN_RPT = 16, SCAN_CH = SCAN_T * SCAN_RPT;

__global__ __launch_bounds__(SCAN_T) void scan_sum_kernel(const int64_t* in, int n, int64_t* bsum) {
    using BR = hipcub::BlockReduce<int64_t, SCAN_T>;
    __shared__ typename BR::TempStorage tmp;
    const int64_t base = (int64_t)blockIdx.x * SCAN_CH;
    int64_t v = 0;
#pragma unroll
    for (int j = 0; j < SCAN_RPT; ++j) {
        const int64_t e = base + j * SCAN_T + threadIdx.x;
        if (e < n) v += in[e];
    }
    const int64_t tot = BR(tmp).Sum(v);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_chunks_kernel(int64_t* bsum, int nb) {
    using BS = hipcub::BlockScan<int64_t, 1024>;
    __shared__ typename BS::TempStorage tmp;
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        const int b = b0 + threadIdx.x;
        const int64_t v = b < nb ? bsum[b] : 0;
        int64_t ex, agg;
        BS(tmp).ExclusiveSum(v, ex, agg);
        const int64_t c = carry;
        if (b < nb) bsum[b] = c + ex;
        __syncthreads();                // every thread has read carry (and the scan's storage is free again)
        if (threadIdx.x == 0) carry = c + agg;
        __syncthreads();
    }
}

__global__ __launch_bounds__(SCAN_T) void scan_apply_kernel(const int64_t* in, int n, const int64_t* boff, int64_t* out) {
    using BS = hipcub::BlockScan<int64_t, SCAN_T>;
    __shared__ typename BS::TempStorage tmp;
    __shared__ int64_t sv[SCAN_CH + SCAN_CH / SCAN_RPT];
    const int64_t base = (int64_t)blockIdx.x * SCAN_CH;
#pragma unroll
    for (int j = 0; j < SCAN_RPT; ++j) {
        const int e = j * SCAN_T + threadIdx.x;
        sv[e + e / SCAN_RPT] = base + e < n ? in[base + e] : 0;
    }
    __syncthreads();
    // the thread's run: elements 16 t .. 16 t + 15 at 17 t .. 17 t + 15
    const int r0 = threadIdx.x * (SCAN_RPT + 1);
    int64_t run[SCAN_RPT];
    int64_t acc = 0;
#pragma unroll
    for (int q = 0; q < SCAN_RPT; ++q) {
        acc += sv[r0 + q];
        run[q] = acc;
    }
    int64_t ex;
    BS(tmp).ExclusiveSum(acc, ex);
    const int64_t off = boff[blockIdx.x] + ex;
#pragma unroll
    for (int q = 0; q < SCAN_RPT; ++q) sv[r0 + q] = off + run[q];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_RPT; ++j) {
        const int e = j * SCAN_T + threadIdx.x;
        if (base + e < n) out[base + e] = sv[e + e / SCAN_RPT];
    }
}

// adjacency scratch: pair counts [n] i64 | the three list counts (256 bytes) | wide list [n] i32 | walk list [n] i32 |
// exact list [n] i32 | scan temporary
struct Adj3Layout {
    size_t counts, list_counts, wide_list, walk_list, exact_list, tmp, total;
};
Adj3Layout adj3_layout(int n) {
    Adj3Layout L{};
    size_t tmp = 0;
    (void)hipcub::DeviceScan::InclusiveSum(nullptr, tmp, (const int64_t*)nullptr, (int64_t*)nullptr, n);
    tmp = std::max(tmp, sizeof(int64_t) * (size_t)((n + SCAN_CH - 1) / SCAN_CH));   // or the chunk sums
    L.counts = 0;
    L.list_counts = al256(sizeof(int64_t) * (size_t)n);
    L.wide_list = L.list_counts + 256;
    L.walk_list = al256(L.wide_list + sizeof(int32_t) * (size_t)n);
    L.exact_list = al256(L.walk_list + sizeof(int32_t) * (size_t)n);
    L.tmp = al256(L.exact_list + sizeof(int32_t) * (size_t)n);
    L.total = al256(L.tmp + tmp);
    return L;
}

extern "C" size_t wsb_adjacency3d_scratch_bytes(int n) {
    if (n <= 0) return 0;
    return adj3_layout(n).total;
}

static int adjacency3d_grid(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                            int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, double* prow, void* scratch,
                            size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(!m->host_only, WSB_EINVAL, "host-only mesh has no device tables");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    WSB_REQUIRE(offsets, WSB_EINVAL, "null offsets");
    hipStream_t s = wsb::as_stream(stream);
    hipLaunchKernelGGL(offsets_head_kernel, dim3(1), dim3(1), 0, s, offsets);
    WSB_HIP(hipGetLastError());
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(pos && sup, WSB_EINVAL, "null buffer");
    const size_t need = wsb_adjacency3d_scratch_bytes(n);
    WSB_REQUIRE(scratch && scratch_bytes >= need, WSB_ENOMEM, "adjacency needs %zu scratch bytes (got %zu)", need, scratch_bytes);
    const Adj3Layout lay = adj3_layout(n);
    char* sb = static_cast<char*>(scratch);
    int64_t* counts = reinterpret_cast<int64_t*>(sb + lay.counts);
    int32_t* list_counts = reinterpret_cast<int32_t*>(sb + lay.list_counts);     // wide, walk, exact
    int32_t* wide_list = reinterpret_cast<int32_t*>(sb + lay.wide_list);
    int32_t* walk_list = reinterpret_cast<int32_t*>(sb + lay.walk_list);
    int32_t* exact_list = reinterpret_cast<int32_t*>(sb + lay.exact_list);
    char* tmp = sb + lay.tmp;                    // the scan's chunk sums
    const bool split = true;
    Adj3Args a{m->dg, pos, sup, order, n, counts, offsets, faces, capacity, chi, list_counts, wide_list, list_counts + 1,
               walk_list, n < WIDE3_MAX_N ? WIDE3 : WIDE3_LARGE, list_counts + 2, exact_list, false, nullptr, nullptr, nullptr, nullptr,
               reinterpret_cast<double4*>(prow), m->support_max * (1.0 + 1e-12), list_counts + 4};
    const dim3 grid((n + 255) / 256), block(256);
    // up to 8,192 wavefronts (32 per CU) for the wide particles, one per particle at a time
    const dim3 wgrid((unsigned)std::min<int64_t>(((int64_t)n + 3) / 4, 2048));
    hipLaunchKernelGGL(zero3_counts_kernel, dim3(1), dim3(1), 0, s, list_counts);
    hipLaunchKernelGGL(adj3_kernel<0>, grid, block, 0, s, a);
    WSB_HIP(hipGetLastError());
    if (split) {
        hipLaunchKernelGGL(adj3_wide_kernel<false>, wgrid, block, 0, s, a);
        WSB_HIP(hipGetLastError());
    }
    {
        const int nb = (int)(((int64_t)n + SCAN_CH - 1) / SCAN_CH);
        int64_t* bsum = reinterpret_cast<int64_t*>(tmp);
        hipLaunchKernelGGL(scan_sum_kernel, dim3(nb), dim3(SCAN_T), 0, s, counts, n, bsum);
        hipLaunchKernelGGL(scan_chunks_kernel, dim3(1), dim3(1024), 0, s, bsum, nb);
        hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(SCAN_T), 0, s, counts, n, bsum, offsets + 1);
        WSB_HIP(hipGetLastError());
    }
    if (faces && capacity > 0) {
        hipLaunchKernelGGL(adj3_kernel<1>, grid, block, 0, s, a);
        WSB_HIP(hipGetLastError());
        if (split) {
            hipLaunchKernelGGL(adj3_wide_kernel<true>, wgrid, block, 0, s, a);
            WSB_HIP(hipGetLastError());
        }
    }
    if (chi) {
        if (split) {
            hipLaunchKernelGGL(adj3_walk_kernel, wgrid, block, 0, s, a);
            WSB_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(adj3_kernel<2>, grid, block, 0, s, a);
        WSB_HIP(hipGetLastError());
        // rare (particles within rounding of a face plane, rays through edges): 1,024 wavefronts at most
        hipLaunchKernelGGL(adj3_exact_kernel, dim3((unsigned)std::min<int64_t>(((int64_t)n + 3) / 4, 256)), block, 0, s, a);
        WSB_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(adj3_mark_kernel, dim3(1), dim3(1), 0, s, a.bad, offsets + n);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_adjacency3d(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                               int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, void* scratch,
                               size_t scratch_bytes, void* stream) {
    return adjacency3d_grid(m, pos, sup, order, n, offsets, faces, capacity, chi, nullptr, scratch, scratch_bytes, stream);
}

extern "C" int wsb_adjacency3d_sorted(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                                      int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, double* prow,
                                      void* scratch, size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(n <= 0 || (order && chi && prow), WSB_EINVAL, "the sorted layout needs order, chi and prow");
    return adjacency3d_grid(m, pos, sup, order, n, offsets, faces, capacity, chi, prow, scratch, scratch_bytes, stream);
}

extern "C" int wsb_adjacency3d_bvh(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                                   int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, void* scratch,
                                   size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(!m->host_only, WSB_EINVAL, "host-only mesh has no device tables");
    WSB_REQUIRE(m->d_bnode, WSB_EINVAL, "mesh has no face tree");
    WSB_REQUIRE(m->nf <= BVH_BITS, WSB_ENOTSUP, "face-tree adjacency takes meshes of at most %d faces (%d)", BVH_BITS, m->nf);
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    WSB_REQUIRE(offsets, WSB_EINVAL, "null offsets");
    hipStream_t s = wsb::as_stream(stream);
    hipLaunchKernelGGL(offsets_head_kernel, dim3(1), dim3(1), 0, s, offsets);
    WSB_HIP(hipGetLastError());
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(pos && sup, WSB_EINVAL, "null buffer");
    const size_t need = wsb_adjacency3d_scratch_bytes(n);
    WSB_REQUIRE(scratch && scratch_bytes >= need, WSB_ENOMEM, "adjacency needs %zu scratch bytes (got %zu)", need, scratch_bytes);
    const Adj3Layout lay = adj3_layout(n);
    char* sb = static_cast<char*>(scratch);
    int64_t* counts = reinterpret_cast<int64_t*>(sb + lay.counts);
    int32_t* list_counts = reinterpret_cast<int32_t*>(sb + lay.list_counts);     // wide (unused), walk, exact, long
    int32_t* long_list = reinterpret_cast<int32_t*>(sb + lay.wide_list);
    int32_t* walk_list = reinterpret_cast<int32_t*>(sb + lay.walk_list);
    int32_t* exact_list = reinterpret_cast<int32_t*>(sb + lay.exact_list);
    char* tmp = sb + lay.tmp;
    size_t tmp_bytes = scratch_bytes - lay.tmp;
    const bool split = n < WIDE3_MAX_N;
    Adj3Args a{m->dg, pos, sup, order, n, counts, offsets, faces, capacity, chi, list_counts, long_list, list_counts + 1,
               walk_list, split ? WIDE3 : INT_MAX, list_counts + 2, exact_list, true, m->d_bnode, m->d_bitem,
               list_counts + 3, long_list, nullptr, m->support_max * (1.0 + 1e-12), list_counts + 4};
    const dim3 grid((n + 255) / 256), block(256);
    const dim3 wgrid((unsigned)std::min<int64_t>(((int64_t)n + 3) / 4, 2048));
    hipLaunchKernelGGL(zero3_counts_kernel, dim3(1), dim3(1), 0, s, list_counts);
    if (chi) hipLaunchKernelGGL(adj3_kernel<0>, grid, block, 0, s, a);      // winding lookups and walk lists only
    hipLaunchKernelGGL(bvh3_kernel<0>, grid, block, 0, s, a);
    WSB_HIP(hipGetLastError());
    WSB_HIP(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, counts, offsets + 1, n, s));
    if (faces && capacity > 0) {
        hipLaunchKernelGGL(bvh3_kernel<1>, grid, block, 0, s, a);
        hipLaunchKernelGGL(bvh3_sort_kernel, dim3((unsigned)std::min(n, 1024)), block, 0, s, a);
        WSB_HIP(hipGetLastError());
    }
    if (chi) {
        if (split) {
            hipLaunchKernelGGL(adj3_walk_kernel, wgrid, block, 0, s, a);
            WSB_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(adj3_kernel<2>, grid, block, 0, s, a);
        WSB_HIP(hipGetLastError());
        hipLaunchKernelGGL(adj3_exact_kernel, dim3((unsigned)std::min<int64_t>(((int64_t)n + 3) / 4, 256)), block, 0, s, a);
        WSB_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(adj3_mark_kernel, dim3(1), dim3(1), 0, s, a.bad, offsets + n);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" size_t wsb_moments3d_scratch_bytes(int n) {
    if (n <= 0) return 0;
    return 256 + (((size_t)n * sizeof(int32_t) + 255) & ~(size_t)255);
}

extern "C" int wsb_moments3d(const wsb_mesh3d* m, int kernel_id, int form, const int64_t* offsets, const int32_t* faces,
                             int64_t capacity, const double* chi, const double* pos, const double* sup, int n, int flags, double* out,
                             void* scratch, size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(!m->host_only, WSB_EINVAL, "host-only mesh has no device tables");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(offsets && pos && sup && out, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(chi || (flags & WSB_OUT_NO_INDICATOR), WSB_EINVAL, "winding numbers required");
    WSB_REQUIRE(scratch && scratch_bytes >= wsb_moments3d_scratch_bytes(n), WSB_ENOMEM,
                "3-D moments need %zu scratch bytes (got %zu)", wsb_moments3d_scratch_bytes(n), scratch_bytes);
    int32_t* hc = static_cast<int32_t*>(scratch);
    int32_t* hl = reinterpret_cast<int32_t*>(static_cast<char*>(scratch) + 256);
    Mom3Args a{offsets, faces, faces ? capacity : 0, chi, pos, sup, n, m->d_frec, m->nf, flags, out, hc, hl};
    hipStream_t s = wsb::as_stream(stream);
    switch (form) {
        case WSB_FORM_FP64: return dispatch22<m3::FP64>(kernel_id, a, s);
        case WSB_FORM_MIXED: return dispatch22<m3::MIXED>(kernel_id, a, s);
    }
    return wsb::set_error(WSB_ENOTSUP, "3-D moments support the fp64 and mixed forms (form %d)", form);
}

extern "C" int wsb_moments3d_sorted(const wsb_mesh3d* m, int kernel_id, int form, const int32_t* order, const double* prow,
                                    const int64_t* offsets, const int32_t* faces, int64_t capacity, const double* chi,
                                    int n, int flags, double* out, void* scratch, size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(!m->host_only, WSB_EINVAL, "host-only mesh has no device tables");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(order && prow && offsets && out, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(chi || (flags & WSB_OUT_NO_INDICATOR), WSB_EINVAL, "winding numbers required");
    WSB_REQUIRE(scratch && scratch_bytes >= wsb_moments3d_scratch_bytes(n), WSB_ENOMEM,
                "3-D moments need %zu scratch bytes (got %zu)", wsb_moments3d_scratch_bytes(n), scratch_bytes);
    int32_t* hc = static_cast<int32_t*>(scratch);
    int32_t* hl = reinterpret_cast<int32_t*>(static_cast<char*>(scratch) + 256);
    Mom3Args a{offsets, faces, faces ? capacity : 0, chi, nullptr, nullptr, n, m->d_frec, m->nf, flags, out, hc, hl, 0,
               reinterpret_cast<const double4*>(prow), order};
    hipStream_t s = wsb::as_stream(stream);
    switch (form) {
        case WSB_FORM_FP64: return dispatch22<m3::FP64>(kernel_id, a, s);
        case WSB_FORM_MIXED: return dispatch22<m3::MIXED>(kernel_id, a, s);
    }
    return wsb::set_error(WSB_ENOTSUP, "3-D moments support the fp64 and mixed forms (form %d)", form);
}

extern "C" int wsb_hessian3d(const wsb_mesh3d* m, int kernel_id, int form, const int64_t* offsets, const int32_t* faces,
                             int64_t capacity, const double* pos, const double* sup, int n, int flags, double* out,
                             void* scratch, size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(!m->host_only, WSB_EINVAL, "host-only mesh has no device tables");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(offsets && pos && sup && out, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(!(flags & ~(WSB_OUT_PHYSICAL | WSB_OUT_ACCUMULATE)), WSB_EINVAL,
                "3-D Hessian flags: WSB_OUT_PHYSICAL, WSB_OUT_ACCUMULATE (got %d)", flags);
    WSB_REQUIRE(scratch && scratch_bytes >= wsb_moments3d_scratch_bytes(n), WSB_ENOMEM,
                "3-D Hessian needs %zu scratch bytes (got %zu)", wsb_moments3d_scratch_bytes(n), scratch_bytes);
    int32_t* hc = static_cast<int32_t*>(scratch);
    int32_t* hl = reinterpret_cast<int32_t*>(static_cast<char*>(scratch) + 256);
    Mom3Args a{offsets, faces, faces ? capacity : 0, nullptr, pos, sup, n, m->d_frec, m->nf, flags, out, hc, hl};
    hipStream_t s = wsb::as_stream(stream);
    switch (form) {
        case WSB_FORM_FP64: return dispatch22<m3::FP64, 1>(kernel_id, a, s);
        case WSB_FORM_MIXED: return dispatch22<m3::MIXED, 1>(kernel_id, a, s);
    }
    return wsb::set_error(WSB_ENOTSUP, "3-D Hessian supports the fp64 and mixed forms (form %d)", form);
}

// ------------------------------------------------------------------------------------------------ no allocation
// The heavy-row list of wsb_moments3d_static lives in a file-scope device array: its count in the first 256 bytes,
// then one row index per particle.  Each GPU holds its own copy, at its own address, until the process ends; the host
// takes that address once per device (hipGetSymbolAddress with the device current) and passes it to the kernels as
// the caller-supplied buffer of wsb_moments3d.
namespace {

constexpr int64_t kStaticRows = int64_t(1) << 21;             // 2,097,152 particles: 8 MiB + 256 bytes per GPU
constexpr int kStaticHead = 256 / (int)sizeof(int32_t);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) int32_t scratch[kStaticHead + kStaticRows];

std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device` (looked up once per device)
int static_scratch(int device, void** ptr) {
    WSB_REQUIRE(device >= 0 && device < kMaxDevices, WSB_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
        int current = 0;
        WSB_HIP(hipGetDevice(&current));
        if (current != device) WSB_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
        if (e != hipSuccess)
            return wsb::set_error(WSB_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                                  hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return WSB_OK;
}

}  // namespace

extern "C" int64_t wsb_moments3d_static_max_rows(void) { return kStaticRows; }

extern "C" int wsb_moments3d_static(const wsb_mesh3d* m, int kernel_id, int form, const int64_t* offsets,
                                    const int32_t* faces, int64_t capacity, const double* chi, const double* pos,
                                    const double* sup, int n, int flags, double* out, void* stream) {
    WSB_REQUIRE(m, WSB_EINVAL, "null mesh");
    WSB_REQUIRE(n >= 0 && n <= kStaticRows, WSB_EINVAL,
                "wsb_moments3d_static serves at most %lld particles per call (got %d); use wsb_moments3d with a "
                "scratch buffer", (long long)kStaticRows, n);
    if (n == 0) return WSB_OK;
    void* ptr = nullptr;
    const int rc = static_scratch(m->device, &ptr);
    if (rc != WSB_OK) return rc;
    return wsb_moments3d(m, kernel_id, form, offsets, faces, capacity, chi, pos, sup, n, flags, out, ptr,
                         sizeof(int32_t) * (size_t)(kStaticHead + kStaticRows), stream);
}

// ------------------------------------------------------------------------------------------------ host evaluation
extern "C" int wsb_adjacency3d_host(const wsb_mesh3d* m, const double* pos, const double* sup, int n, int64_t* offsets,
                                    int32_t* faces, int64_t capacity, double* chi) {
    WSB_REQUIRE(m && offsets, WSB_EINVAL, "null argument");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    offsets[0] = 0;
    const m3::Grid3& g = m->hg;
    for (int i = 0; i < n; ++i) {
        const double px = pos[3 * (int64_t)i], py = pos[3 * (int64_t)i + 1], pz = pos[3 * (int64_t)i + 2], h = sup[i];
        int c[3];
        int64_t cnt = 0;
        double w = 0.0;
        if (m3::cell_of(g, px, py, pz, c)) {
            const int64_t cid = m3::cell_id(g, c[0], c[1], c[2]);
            for (int it = g.cstart[cid]; it < g.cstart[cid + 1]; ++it) {
                const int f = g.citems[it];
                const double* bx = g.fbox + 6 * (int64_t)f;
                if (box_far(bx, px, py, pz, h * (1.0 + 1e-9)) || !keep_face(bx, px, py, pz, h)) continue;
                if (faces && offsets[i] + cnt < capacity) faces[offsets[i] + cnt] = f;
                ++cnt;
            }
            if (!m->soup && !m3::ray_winding(g, px, py, pz, c, &w)) w = exact_winding(g.frec, g.nf, px, py, pz, 1.0 / h);
        }
        offsets[i + 1] = offsets[i] + cnt;
        if (chi) chi[i] = w;
    }
    return WSB_OK;
}

extern "C" int wsb_winding3d_exact_host(const wsb_mesh3d* m, const double* pos, int n, double* chi) {
    WSB_REQUIRE(m && pos && chi, WSB_EINVAL, "null argument");
    for (int i = 0; i < n; ++i)
        chi[i] = exact_winding(m->hg.frec, m->nf, pos[3 * (int64_t)i], pos[3 * (int64_t)i + 1], pos[3 * (int64_t)i + 2], 1.0);
    return WSB_OK;
}

template <int FORM>
int host_pairs_form(int kernel_id, const double* frec, const int32_t* pf, const double* pq, const double* ph, int P, double* out) {
    switch (kernel_id) {
        case WSB_CUBIC: host_pairs<WSB_CUBIC, FORM>(frec, pf, pq, ph, P, out); return WSB_OK;
        case WSB_W2: host_pairs<WSB_W2, FORM>(frec, pf, pq, ph, P, out); return WSB_OK;
        case WSB_W4: host_pairs<WSB_W4, FORM>(frec, pf, pq, ph, P, out); return WSB_OK;
        case WSB_W6: host_pairs<WSB_W6, FORM>(frec, pf, pq, ph, P, out); return WSB_OK;
    }
    return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
}

extern "C" int wsb_face_terms3d_host(const wsb_mesh3d* m, int kernel_id, int form, const int32_t* pair_face,
                                     const double* pair_pos, const double* pair_sup, int P, double* out) {
    WSB_REQUIRE(m && pair_face && pair_pos && pair_sup && out, WSB_EINVAL, "null argument");
    for (int j = 0; j < P; ++j) WSB_REQUIRE(pair_face[j] >= 0 && pair_face[j] < m->nf, WSB_EINVAL, "face index out of range");
    if (form == WSB_FORM_FP64) return host_pairs_form<m3::FP64>(kernel_id, m->hg.frec, pair_face, pair_pos, pair_sup, P, out);
    if (form == WSB_FORM_MIXED) return host_pairs_form<m3::MIXED>(kernel_id, m->hg.frec, pair_face, pair_pos, pair_sup, P, out);
    return wsb::set_error(WSB_ENOTSUP, "3-D face terms support the fp64 and mixed forms (form %d)", form);
}

extern "C" int wsb_face_hessian3d_host(const wsb_mesh3d* m, int kernel_id, int form, const int32_t* pair_face,
                                       const double* pair_pos, const double* pair_sup, int P, double* out) {
    WSB_REQUIRE(m && pair_face && pair_pos && pair_sup && out, WSB_EINVAL, "null argument");
    for (int j = 0; j < P; ++j) WSB_REQUIRE(pair_face[j] >= 0 && pair_face[j] < m->nf, WSB_EINVAL, "face index out of range");
    WSB_REQUIRE(form == WSB_FORM_FP64 || form == WSB_FORM_MIXED, WSB_ENOTSUP, "3-D Hessian supports the fp64 and mixed forms (form %d)", form);
    const double* fr = m->hg.frec;
    const bool mixed = form == WSB_FORM_MIXED;
    switch (kernel_id) {
        case WSB_CUBIC: mixed ? host_hessian_pairs<WSB_CUBIC, m3::MIXED>(fr, pair_face, pair_pos, pair_sup, P, out)
                              : host_hessian_pairs<WSB_CUBIC, m3::FP64>(fr, pair_face, pair_pos, pair_sup, P, out); return WSB_OK;
        case WSB_W2: mixed ? host_hessian_pairs<WSB_W2, m3::MIXED>(fr, pair_face, pair_pos, pair_sup, P, out)
                           : host_hessian_pairs<WSB_W2, m3::FP64>(fr, pair_face, pair_pos, pair_sup, P, out); return WSB_OK;
        case WSB_W4: mixed ? host_hessian_pairs<WSB_W4, m3::MIXED>(fr, pair_face, pair_pos, pair_sup, P, out)
                           : host_hessian_pairs<WSB_W4, m3::FP64>(fr, pair_face, pair_pos, pair_sup, P, out); return WSB_OK;
        case WSB_W6: mixed ? host_hessian_pairs<WSB_W6, m3::MIXED>(fr, pair_face, pair_pos, pair_sup, P, out)
                           : host_hessian_pairs<WSB_W6, m3::FP64>(fr, pair_face, pair_pos, pair_sup, P, out); return WSB_OK;
    }
    return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
}

extern "C" int wsb_finalize22_host(int kernel_id, const double* acc, const double* chi, const double* sup, int n, int flags,
                                   double* out) {
    WSB_REQUIRE(acc && chi && sup && out, WSB_EINVAL, "null argument");
    switch (kernel_id) {
        case WSB_CUBIC: host_finalize<WSB_CUBIC>(acc, chi, sup, n, flags, out); return WSB_OK;
        case WSB_W2: host_finalize<WSB_W2>(acc, chi, sup, n, flags, out); return WSB_OK;
        case WSB_W4: host_finalize<WSB_W4>(acc, chi, sup, n, flags, out); return WSB_OK;
        case WSB_W6: host_finalize<WSB_W6>(acc, chi, sup, n, flags, out); return WSB_OK;
    }
    return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
}
