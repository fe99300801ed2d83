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
