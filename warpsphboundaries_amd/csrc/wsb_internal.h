// Internal declarations of the wall-integral library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/wsb.h"

namespace wsb {

int set_error(int code, const char* fmt, ...);

#define WSB_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ::wsb::set_error(WSB_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define WSB_REQUIRE(cond, code, ...)                                                           \
    do {                                                                                       \
        if (!(cond)) return ::wsb::set_error((code), __VA_ARGS__);                            \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// the twelve channels: 0 lam 1 m10 2 m01 3 Gx 4 Gy 5 g1xx 6 g1xy 7 g1yx 8 g1yy 9 m20 10 m11 11 m02
constexpr int NCH = 12;

}  // namespace wsb

// separately rounded a*b - c*d (the orientation predicate shared by the indicator and the edge terms;
// reference edge/predicates.py:21-32): no fused multiply-add may merge the products.
__host__ __device__ inline double wsb_cross2(double ax, double ay, double bx, double by) {
#pragma clang fp contract(off)
    double l = ax * by;
    double r = ay * bx;
    return l - r;
}

struct wsb_geom2d {
    int device = 0;
    int nv = 0, ne = 0, background = 0;
    double support_max = 0;
    // static cell list (cell = support_max) and y-bins
    double clo0 = 0, clo1 = 0, cell = 0;
    int nx = 0, ny = 0, K = 1;
    double y0 = 0, dy = 0;
    int nb = 1;
    int64_t ncell_items = 0, nbin_items = 0, bytes = 0;
    // device arrays
    double* verts = nullptr;      // [nv*2]
    int32_t* edges = nullptr;     // [ne*2]
    int32_t* cstart = nullptr;    // [nx*ny+1]
    int32_t* citems = nullptr;    // [ncell_items]
    double* cind = nullptr;       // [nx*ny] indicator of edge-free cells
    int32_t* cempty = nullptr;    // [nx*ny]
    int32_t* bstart = nullptr;    // [nb+1]
    int32_t* bitems = nullptr;    // [nbin_items]
    // coordinate records (one load instead of edge -> vertex indirection)
    double4* cxy = nullptr;       // [ncell_items] (ax, ay, bx, by) of each cell item
    double4* bxy = nullptr;       // [nbin_items]  (ax, ay, bx, by) of each bin item
    double4* erec = nullptr;      // [ne] (ax, ay, bx, by)
    double4* etan = nullptr;      // [ne] (tx, ty, 1/len, len): unit tangent and length in the local frame
    // local winding: per cell with edges a reference point and its indicator (x, y, indicator, 1), (0, 0, 0, 0)
    // where none qualifies; `local` = 0 when some edge is too close to horizontal for the ray argument
    double4* cref = nullptr;      // [nx*ny]
    // edge tree of the nearest-edge search: node boxes, links (inner: right child, left = next node; leaf:
    // (-(first + 1), count)) and the edge ids its leaves range over; extent = the largest |coordinate| of a vertex
    double4* tbox = nullptr;      // [tnodes]
    int2* tlink = nullptr;        // [tnodes]
    int32_t* titems = nullptr;    // [ne]
    int tnodes = 0;
    double extent = 0;
    int local = 0;
    int64_t noref = 0;            // cells with edges and no reference point
    // nearest-edge candidate grid (built on the device by the first signed-distance call): ng x ng cells over the
    // vertex box widened by its larger side on every side; per cell the edges that can be nearest to a point of the
    // cell (CSR, edge ids); an empty list sends the cell's points to the edge tree
    int ng = 0;
    double nlo0 = 0, nlo1 = 0, ncs = 0;
    int32_t* nstart = nullptr;    // [ng*ng+1]
    int32_t* nitems = nullptr;    // [nstart[ng*ng]]
};

struct wsb_plan {
    int device = 0;
    int kernel_id = 0;
    int nr = 0, ncoef = 0, nE = 0, nV = 0;
    double* radii = nullptr;
    int32_t* cn = nullptr;
    double* cc = nullptr;
    int32_t* E = nullptr;   // [nE*7]
    int32_t* V = nullptr;   // [nV*4]
    double* mR = nullptr;   // [nV]
};
