// Conditioned small / sliver triangle: the twelve channels of a closed 3-vertex loop by an adaptive area quadrature,
// for the rows whose support makes the triangle small (longest edge <= 0.02 h) or a sliver (twice the area <= 0.01
// longest edge squared), where the edge-local closed form sums edge terms much larger than the result.
//
// Algorithm (reference edge/warptriangle.py:27-181): with a = (v0 - x)/h, b = (v1 - v0)/h, c = (v2 - v0)/h, the
// triangle is a + u b + t (1 - u) c, (u, t) in [0, 1]^2, Jacobian (1 - u) |b x c|.  The outer variable u is cut where
// the two long edges' lines pass closest to the particle or cross a support circle, and where the inner segment (the
// direction c) passes through the particle or at a support radius from it; each outer interval is bisected until a
// 16-node Gauss-Legendre panel and an 8-node one (both with 32-node inner rules on the inner segment, itself cut at
// its closest point and circle crossings) agree to `tolerance` (hi - lo + max |panel|) in channels scaled by the
// triangle's extent; the converged panels add the 16-node values, and |16 - 8| + |16 nodes x 32 - 16 nodes x 16|
// to the estimate.  A panel that cannot be split (stack of 62, `max_panels` visited, width 1e-14) marks the row -1
// (values NaN).  The triangle is ordered with its shortest edge first, so the inner segments run across the sliver.
#include <cmath>

#include "edge2d_device.h"
#include "wsb_internal.h"

namespace tri {

constexpr int MAXCUT = 32;
constexpr int STACK = 64;
constexpr double INV_PI = 0.318309886183790671537767526745028724;

struct Tri {
    double v0x, v0y, e0x, e0y, e1x, e1y;   // ordered vertices: v0, e0 = v1 - v0, e1 = v2 - v0
    double area2, sign, background;        // |twice the signed area|, its sign (+1 counter-clockwise), background
};

__host__ __device__ inline void add_cut(double* cuts, int& n, double t) {
    if (t > 0.0 && t < 1.0 && n < MAXCUT) cuts[n++] = t;
}

// cuts of the line a + t b, t in (0, 1): its closest point to the particle and its support-circle crossings
template <int KID>
__host__ __device__ inline void line_cuts(double* cuts, int& n, double ax, double ay, double bx, double by) {
    const double bb = bx * bx + by * by;
    if (!(bb > 0.0)) return;
    const double t = -(ax * bx + ay * by) / bb;
    add_cut(cuts, n, t);
    const double cx = ax + t * bx, cy = ay + t * by;
    const double rr = cx * cx + cy * cy;
    for (int j = 0; j < TTab<KID>::NRAD; ++j) {
        const double R = TTab<KID>::RAD[j];
        const double delta = (R * R - rr) / bb;
        if (delta >= 0.0) {
            const double root = sqrt(delta);
            add_cut(cuts, n, t - root);
            add_cut(cuts, n, t + root);
        }
    }
}

__host__ __device__ inline void sort_cuts(double* cuts, int n) {
    for (int i = 1; i < n; ++i) {
        const double key = cuts[i];
        int j = i - 1;
        while (j >= 0 && cuts[j] > key) {
            cuts[j + 1] = cuts[j];
            --j;
        }
        cuts[j + 1] = key;
    }
}

// pi W(r) and -(pi W)'(r) / r (units of h): the truncated blocks in powers of u = R - r, and near the origin the
// first piece's polynomial, whose -W'/r is a polynomial (W'(0) = 0)
template <int KID>
__host__ __device__ inline void kernel_wr(double r, double& W, double& radial) {
    if constexpr (KID >= WSB_W2P5) {           // W^5: the factored form (its powers of u cancel near r = 0)
        chord::p5_wr<KID>(r, W, radial);
        return;
    }
    using T = TTab<KID>;
    W = 0.0;
    double dw = 0.0;
    for (int j = 0; j < T::NRAD; ++j) {
        if (r < T::RAD[j]) {
            const double z = T::RAD[j] - r;
            double p = 0.0, dp = 0.0;
            for (int k = T::DEG; k >= 1; --k) {
                p = p * z + T::UB[j][k];
                dp = dp * z + k * T::UB[j][k];
            }
            W += p * z + T::UB[j][0];
            dw -= dp;                          // d/dr = -d/dz
        }
    }
    radial = -dw / fmax(r, 1e-300);
    if (T::REGULAR && r < T::ORIG_HI) {
        radial = 0.0;
        for (int k = T::NORIG - 1; k >= 2; --k) radial = radial * r - k * T::ORIG[k];
    }
}

// the inner segment at u: start a + u b, direction (1 - u) c; `count` Gauss nodes per piece between its cuts;
// adds w_u * (1 - u) / pi * the channels / scales to out
template <int KID>
__host__ __device__ inline void line_sum(double u, double wu, const double* a, const double* b, const double* c,
                                         const double* inv_scale, const double* gx, const double* gw, int count, double* out) {
    const double sx = a[0] + u * b[0], sy = a[1] + u * b[1];
    const double dx = (1.0 - u) * c[0], dy = (1.0 - u) * c[1];
    double cuts[MAXCUT];
    cuts[0] = 0.0;
    cuts[1] = 1.0;
    int n = 2;
    line_cuts<KID>(cuts, n, sx, sy, dx, dy);
    sort_cuts(cuts, n);
    const double wl = wu * (1.0 - u) * INV_PI;
    for (int part = 0; part + 1 < n; ++part) {
        const double lo = cuts[part], hi = cuts[part + 1];
        if (!(hi > lo)) continue;
        for (int j = 0; j < count; ++j) {
            const double t = lo + (hi - lo) * gx[j];
            const double y0 = sx + t * dx, y1 = sy + t * dy;
            const double r = sqrt(y0 * y0 + y1 * y1);
            double W, radial;
            kernel_wr<KID>(r, W, radial);
            const double Gx = radial * y0, Gy = radial * y1;
            const double w = (hi - lo) * gw[j] * wl;
            const double v[wsb::NCH] = {W, y0 * W, y1 * W, Gx, Gy, y0 * Gx, y0 * Gy, y1 * Gx, y1 * Gy,
                                        y0 * y0 * W, y0 * y1 * W, y1 * y1 * W};
#pragma unroll
            for (int k = 0; k < wsb::NCH; ++k) out[k] += w * v[k] * inv_scale[k];
        }
    }
}

// one outer panel [lo, hi]: kind 1 = 16 outer x 32 inner nodes, 0 = 8 x 32, -1 = 16 x 16
template <int KID>
__host__ __device__ inline void panel(double lo, double hi, int kind, const double* a, const double* b, const double* c,
                                      const double* inv_scale, double* out) {
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) out[k] = 0.0;
    const double* ox = kind != 0 ? GLTab::X16 : GLTab::X8;
    const double* ow = kind != 0 ? GLTab::W16 : GLTab::W8;
    const int ocount = kind != 0 ? 16 : 8;
    const double* ix = kind < 0 ? GLTab::X16 : GLTab::X32;
    const double* iw = kind < 0 ? GLTab::W16 : GLTab::W32;
    const int icount = kind < 0 ? 16 : 32;
    for (int j = 0; j < ocount; ++j)
        line_sum<KID>(lo + (hi - lo) * ox[j], (hi - lo) * ow[j], a, b, c, inv_scale, ix, iw, icount, out);
}

// one row: status 0 (closed-form route: the triangle is neither small nor a sliver at this support; values and
// estimate untouched), 1 (quadrature converged) or -1 (budget exhausted: values NaN); values / estimate in units of h
template <int KID>
__host__ __device__ inline int triangle_row(const Tri& T, double px, double py, double h, double tol, int max_panels,
                                            double* values, double* estimate) {
    const double l0 = T.e0x * T.e0x + T.e0y * T.e0y, l1 = T.e1x * T.e1x + T.e1y * T.e1y;
    const double fx = T.e1x - T.e0x, fy = T.e1y - T.e0y;
    const double longest2 = fmax(l0, fmax(l1, fx * fx + fy * fy));
    if (longest2 > 0.02 * 0.02 * h * h && T.area2 > 0.01 * longest2) return 0;
    int status = 1;
    const double ih = 1.0 / h;
    const double a[2] = {(T.v0x - px) * ih, (T.v0y - py) * ih};
    const double b[2] = {T.e0x * ih, T.e0y * ih}, c[2] = {T.e1x * ih, T.e1y * ih};
    const double sx = fmax(fmax(fabs(a[0]), fmax(fabs(a[0] + b[0]), fabs(a[0] + c[0]))), 1e-150);
    const double sy = fmax(fmax(fabs(a[1]), fmax(fabs(a[1] + b[1]), fabs(a[1] + c[1]))), 1e-150);
    const double scale[wsb::NCH] = {1.0, sx, sy, sx, sy, sx * sx, sx * sy, sy * sx, sy * sy, sx * sx, sx * sy, sy * sy};
    double inv_scale[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) inv_scale[k] = 1.0 / scale[k];
    // outer cuts: the edge v0 -> v1 (a + u b) and the edge v2 -> v1 ((a + c) + u (b - c)); where the inner segment
    // (direction c from a + u b) passes through the particle and at a support radius from it
    double cuts[MAXCUT];
    cuts[0] = 0.0;
    cuts[1] = 1.0;
    int n = 2;
    line_cuts<KID>(cuts, n, a[0], a[1], b[0], b[1]);
    line_cuts<KID>(cuts, n, a[0] + c[0], a[1] + c[1], b[0] - c[0], b[1] - c[1]);
    const double bc = b[0] * c[1] - b[1] * c[0];
    const double ac = a[0] * c[1] - a[1] * c[0];
    if (bc != 0.0) {
        add_cut(cuts, n, -ac / bc);
        const double lc = sqrt(c[0] * c[0] + c[1] * c[1]);
        for (int j = 0; j < TTab<KID>::NRAD; ++j)
            for (int s = -1; s <= 1; s += 2) add_cut(cuts, n, (s * TTab<KID>::RAD[j] * lc - ac) / bc);
    }
    sort_cuts(cuts, n);
    double total[wsb::NCH], errors[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) total[k] = errors[k] = 0.0;
    int visited = 0;
    double slo[STACK], shi[STACK];
    for (int part = 0; part + 1 < n; ++part) {
        slo[0] = cuts[part];
        shi[0] = cuts[part + 1];
        int size = 1;
        while (size > 0) {
            --size;
            const double lo = slo[size], hi = shi[size];
            if (!(hi > lo)) continue;
            double fine[wsb::NCH], coarse[wsb::NCH];
            panel<KID>(lo, hi, 1, a, b, c, inv_scale, fine);
            panel<KID>(lo, hi, 0, a, b, c, inv_scale, coarse);
            double err = 0.0, norm = 0.0;
#pragma unroll
            for (int k = 0; k < wsb::NCH; ++k) {
                err = fmax(err, fabs(fine[k] - coarse[k]));
                norm = fmax(norm, fabs(fine[k]));
            }
            ++visited;
            const bool accept = err <= tol * (hi - lo + norm);
            if (!accept && size < STACK - 2 && visited < max_panels && hi - lo > 1e-14) {
                const double mid = 0.5 * (lo + hi);
                slo[size] = mid;
                shi[size] = hi;
                slo[size + 1] = lo;
                shi[size + 1] = mid;
                size += 2;
            } else {
                if (!accept) status = -1;
                double inner[wsb::NCH];
                panel<KID>(lo, hi, -1, a, b, c, inv_scale, inner);
#pragma unroll
                for (int k = 0; k < wsb::NCH; ++k) {
                    total[k] += fine[k];
                    errors[k] += fabs(fine[k] - coarse[k]) + fabs(fine[k] - inner[k]);
                }
            }
        }
    }
    const double jac = T.area2 * ih * ih;
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) {
        values[k] = status < 0 ? __builtin_nan("") : T.background * TTab<KID>::PLANE[k] + T.sign * total[k] * scale[k] * jac;
        estimate[k] = errors[k] * scale[k] * jac;
    }
    return status;
}

struct TriArgs {
    Tri T;
    const double* lpos;
    const double* lsup;
    const uint8_t* valid;
    int n;
    double tol;
    int max_panels;
    double* values;
    double* estimate;
    int32_t* status;
    const double* rot4;
    int flags;
    double* out;
};

// one thread per row; rows that take the quadrature also overwrite their row of `out` (finalised with `flags`)
template <int KID>
__global__ __launch_bounds__(64) void triangle_kernel(TriArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double v[wsb::NCH], e[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) v[k] = e[k] = 0.0;
    int st = 0;
    const bool valid = !a.valid || a.valid[i];
    const double h = a.lsup[i];
    if (valid) st = triangle_row<KID>(a.T, a.lpos[2 * (int64_t)i], a.lpos[2 * (int64_t)i + 1], h, a.tol, a.max_panels, v, e);
    if (st == 0) {
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) v[k] = e[k] = 0.0;
    }
    if (a.values) {
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) a.values[(int64_t)i * wsb::NCH + k] = v[k];
    }
    if (a.estimate) {
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) a.estimate[(int64_t)i * wsb::NCH + k] = e[k];
    }
    if (a.status) a.status[i] = st;
    if (st != 0 && a.out)
        finalize_row(v, 0.0, h, TTab<KID>::PLANE, a.rot4, a.flags | WSB_OUT_NO_INDICATOR, a.out + (int64_t)i * wsb::NCH);
}

template <int KID>
void host_rows(const TriArgs& a) {
    for (int i = 0; i < a.n; ++i) {
        double v[wsb::NCH] = {}, e[wsb::NCH] = {};
        int st = 0;
        if (!a.valid || a.valid[i])
            st = triangle_row<KID>(a.T, a.lpos[2 * (int64_t)i], a.lpos[2 * (int64_t)i + 1], a.lsup[i], a.tol, a.max_panels, v, e);
        if (st == 0)
            for (int k = 0; k < wsb::NCH; ++k) v[k] = e[k] = 0.0;
        for (int k = 0; k < wsb::NCH; ++k) {
            if (a.values) a.values[(int64_t)i * wsb::NCH + k] = v[k];
            if (a.estimate) a.estimate[(int64_t)i * wsb::NCH + k] = e[k];
        }
        if (a.status) a.status[i] = st;
    }
}

int make_args(const double* tri, int background, const double* lpos, const double* lsup, const uint8_t* valid, int n,
              double tolerance, int max_panels, double* values, double* estimate, int32_t* status, const double* rot4,
              int flags, double* out, TriArgs* a) {
    WSB_REQUIRE(tri, WSB_EINVAL, "null triangle");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    WSB_REQUIRE(tolerance > 0.0 && max_panels >= 1, WSB_EINVAL, "positive tolerance and max_panels required");
    WSB_REQUIRE(!(flags & WSB_OUT_ACCUMULATE), WSB_EINVAL,
                "the conditioned rows overwrite their output rows: pass this representation's own contribution");
    WSB_REQUIRE(n == 0 || (lpos && lsup), WSB_EINVAL, "null buffer");
    for (int k = 0; k < 8; ++k) WSB_REQUIRE(std::isfinite(tri[k]), WSB_EINVAL, "non-finite triangle record");
    WSB_REQUIRE(tri[6] > 0.0 && (tri[7] == 1.0 || tri[7] == -1.0), WSB_EINVAL, "degenerate triangle record");
    a->T = Tri{tri[0], tri[1], tri[2], tri[3], tri[4], tri[5], tri[6], tri[7], (double)background};
    a->lpos = lpos;
    a->lsup = lsup;
    a->valid = valid;
    a->n = n;
    a->tol = tolerance;
    a->max_panels = max_panels;
    a->values = values;
    a->estimate = estimate;
    a->status = status;
    a->rot4 = rot4;
    a->flags = flags;
    a->out = out;
    return WSB_OK;
}

}  // namespace tri

extern "C" int wsb_triangle2d(const double* tri, int background, int kernel_id, const double* lpos, const double* lsup,
                              const uint8_t* valid, int n, double tolerance, int max_panels, double* values, double* estimate,
                              int32_t* status, const double* rot4, int flags, double* out, void* stream) {
    tri::TriArgs a;
    const int st = tri::make_args(tri, background, lpos, lsup, valid, n, tolerance, max_panels, values, estimate, status,
                                  rot4, flags, out, &a);
    if (st) return st;
    if (n == 0) return WSB_OK;
    hipStream_t s = wsb::as_stream(stream);
    const dim3 g((n + 63) / 64), b(64);
    switch (kernel_id) {
        case WSB_CUBIC: hipLaunchKernelGGL(tri::triangle_kernel<WSB_CUBIC>, g, b, 0, s, a); break;
        case WSB_W2: hipLaunchKernelGGL(tri::triangle_kernel<WSB_W2>, g, b, 0, s, a); break;
        case WSB_W4: hipLaunchKernelGGL(tri::triangle_kernel<WSB_W4>, g, b, 0, s, a); break;
        case WSB_W6: hipLaunchKernelGGL(tri::triangle_kernel<WSB_W6>, g, b, 0, s, a); break;
        case WSB_CONE: hipLaunchKernelGGL(tri::triangle_kernel<WSB_CONE>, g, b, 0, s, a); break;
        case WSB_LW2: hipLaunchKernelGGL(tri::triangle_kernel<WSB_LW2>, g, b, 0, s, a); break;
        case WSB_LW4: hipLaunchKernelGGL(tri::triangle_kernel<WSB_LW4>, g, b, 0, s, a); break;
        case WSB_W2P5: hipLaunchKernelGGL(tri::triangle_kernel<WSB_W2P5>, g, b, 0, s, a); break;
        case WSB_W4P5: hipLaunchKernelGGL(tri::triangle_kernel<WSB_W4P5>, g, b, 0, s, a); break;
        default: return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_triangle2d_host(const double* tri, int background, int kernel_id, const double* lpos, const double* lsup,
                                   const uint8_t* valid, int n, double tolerance, int max_panels, double* values,
                                   double* estimate, int32_t* status) {
    tri::TriArgs a;
    const int st = tri::make_args(tri, background, lpos, lsup, valid, n, tolerance, max_panels, values, estimate, status,
                                  nullptr, 0, nullptr, &a);
    if (st) return st;
    switch (kernel_id) {
        case WSB_CUBIC: tri::host_rows<WSB_CUBIC>(a); break;
        case WSB_W2: tri::host_rows<WSB_W2>(a); break;
        case WSB_W4: tri::host_rows<WSB_W4>(a); break;
        case WSB_W6: tri::host_rows<WSB_W6>(a); break;
        case WSB_CONE: tri::host_rows<WSB_CONE>(a); break;
        case WSB_LW2: tri::host_rows<WSB_LW2>(a); break;
        case WSB_LW4: tri::host_rows<WSB_LW4>(a); break;
        case WSB_W2P5: tri::host_rows<WSB_W2P5>(a); break;
        case WSB_W4P5: tri::host_rows<WSB_W4P5>(a); break;
        default: return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
    }
    return WSB_OK;
}
