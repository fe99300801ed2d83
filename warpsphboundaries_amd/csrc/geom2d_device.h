// Device and host arithmetic of the 2-D static geometry shared by the per-scene and the batched kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "wsb_internal.h"

// winding number + background at one point (reference scene/scene.py:333-371 and scene/fixedadj.py:98-146);
// `h` normalises the relative coordinates (h = 0: unnormalised, the cell-centre convention).
//
// The reference tests the sign of the separately rounded cross product of the support-normalised relative
// coordinates.  Each bin edge is first tested on the unnormalised differences: when |cu| exceeds 16 u times the
// magnitude of its two products (u = 2^-53) the exact cross product of the rounded differences, and hence the
// normalised rounded one, has the sign of cu and is nonzero (rounding of the products, the subtraction and the
// divisions by h each perturb it by at most 5 u of that magnitude).  Otherwise the reference's own arithmetic is
// evaluated.  Edges wholly left of the point neither cross the ray nor contain the point: their cross product is
// strictly negative for an upward and strictly positive for a downward edge, so they change nothing.
__host__ __device__ inline double winding_point(double px, double py, double h, const double* V, const int32_t* E, int ne,
                                                const int32_t* bstart, const double4* bxy, double y0, double dy, int nb,
                                                double bg) {
#pragma clang fp contract(off)
    const double iyf = floor((py - y0) / dy);
    if (!(iyf >= 0.0 && iyf <= (double)nb)) return bg;
    const int iy = (int)iyf < nb - 1 ? (int)iyf : nb - 1;
    double w = 0.0;
    bool boundary = false;
    for (int i = bstart[iy]; i < bstart[iy + 1]; ++i) {
        const double4 rec = bxy[i];
        const double ax = rec.x, ay = rec.y;
        const double cxx = rec.z, cyy = rec.w;
        if (ax < px && cxx < px) continue;
        const bool up = ay <= py && py < cyy;
        const bool down = cyy <= py && py < ay;
        const double ux = ax - px, uy = ay - py, vx = cxx - px, vy = cyy - py;
        const double l = ux * vy, r = uy * vx;
        const double mag = fabs(l) + fabs(r);
        const double cu = l - r;
        if (h > 0.0 && mag > 1e-280 && fabs(cu) > 1.7763568394002505e-15 * mag) {   // 16 u
            if (up && cu > 0.0) w += 1.0;
            if (down && cu < 0.0) w -= 1.0;
            continue;
        }
        double arx, ary, crx, cry;
        if (h > 0.0) {
            arx = ux / h;
            ary = uy / h;
            crx = vx / h;
            cry = vy / h;
        } else {
            arx = ux;
            ary = uy;
            crx = vx;
            cry = vy;
        }
        const double cross = wsb_cross2(arx, ary, crx, cry);
        if (up && cross > 0.0) w += 1.0;
        if (down && cross < 0.0) w -= 1.0;
        if (cross == 0.0 && arx * crx + ary * cry <= 0.0) boundary = true;
    }
    if (boundary) {
        double angle = 0.0;
        for (int e = 0; e < ne; ++e) {
            const int v0 = E[2 * e], v1 = E[2 * e + 1];
            double arx, ary, crx, cry;
            if (h > 0.0) {
                arx = (V[2 * v0] - px) / h;
                ary = (V[2 * v0 + 1] - py) / h;
                crx = (V[2 * v1] - px) / h;
                cry = (V[2 * v1 + 1] - py) / h;
            } else {
                arx = V[2 * v0] - px;
                ary = V[2 * v0 + 1] - py;
                crx = V[2 * v1] - px;
                cry = V[2 * v1 + 1] - py;
            }
            const double cross = wsb_cross2(arx, ary, crx, cry);
            if (cross != 0.0) angle += atan2(cross, arx * crx + ary * cry);
        }
        w = angle / 6.283185307179586;
    }
    return w + bg;
}

// Sign of the orientation (a, b, c) = (a - c) x (b - c) when the rounded value certainly has the sign of the exact
// one (|value| above 16 u of the magnitude of its two products, the filter of winding_point), else 0.  With c = p
// it is the ray test's cu of winding_point.
__host__ __device__ inline int orient_certain(double ax, double ay, double bx, double by, double cx, double cy) {
#pragma clang fp contract(off)
    const double l = (ax - cx) * (by - cy), r = (ay - cy) * (bx - cx);
    const double mag = fabs(l) + fabs(r), d = l - r;
    if (!(mag > 1e-280 && fabs(d) > 1.7763568394002505e-15 * mag)) return 0;
    return d > 0.0 ? 1 : -1;
}

// One edge (a, b) of a point's cell list in the local winding count: the winding number at p is the one at the
// cell's reference point q plus the signed crossings of the segment q -> p (+1 from the right of a -> b to its
// left).  Every edge that can cross the segment lies in the cell list.  unsure is set when the reference's own ray
// decision at p for this edge is uncertain (p within rounding of the edge's line while the edge is not wholly left
// of p: there the reference may differ from the topology) or when a crossing decision is uncertain; the caller
// then evaluates winding_point.  Edges outside the cell list are never uncertain for the reference when every
// non-horizontal edge of the geometry is steep enough (checked once per geometry, wsb_geom2d_create).
__host__ __device__ inline void local_winding_edge(double ax, double ay, double bx, double by, double px, double py,
                                                   double qx, double qy, int& dw, bool& unsure) {
    const int sp = orient_certain(ax, ay, bx, by, px, py);
    if (sp == 0 && !(ax < px && bx < px)) unsure = true;
    if (fmax(ax, bx) < fmin(qx, px) || fmin(ax, bx) > fmax(qx, px) || fmax(ay, by) < fmin(qy, py) ||
        fmin(ay, by) > fmax(qy, py))
        return;
    const int sq = orient_certain(ax, ay, bx, by, qx, qy);
    const int sa = orient_certain(qx, qy, px, py, ax, ay);
    const int sb = orient_certain(qx, qy, px, py, bx, by);
    if (sq * sp < 0 && sa * sb < 0) dw += sp;
    else if (!(sq * sp > 0 || sa * sb > 0)) unsure = true;
}

