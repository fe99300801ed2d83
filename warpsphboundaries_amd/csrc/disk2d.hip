// Solid disks of a body frame (a fibre bundle, or one solid disk): the static cell list over the disk centres, the
// fixed-capacity slot fill of every particle, and the twelve channels of every (particle, disk) pair, FP64.
//
// Channels of one disk (h = 1, centre at distance D along the unit vector e from the particle, radius R, tangent t):
// with y the offset from the particle, r = |y| and alpha(r) the half angle of the arc of the circle |y| = r inside the
// disk (pi where the whole circle is inside, 0 where none of it is), the polar integrals reduce to one radial
// integral each:
//     lam  = int W                =  int_0^1 W(r) 2 alpha r dr
//     m1   = int y W              =  e int_0^1 W(r) 2 sin(alpha) r^2 dr
//     g0   = int grad_x W         = -e int_0^1 W'(r) 2 sin(alpha) r dr
//     g1   = int y (x) grad_x W   = -int_0^1 W'(r) r^2 [(alpha + sin cos) e e^T + (alpha - sin cos) t t^T] dr
//     m2   = int y (x) y W        =  int_0^1 W(r) r^3 [(alpha + sin cos) e e^T + (alpha - sin cos) t t^T] dr
// cos(alpha) = (r^2 + D^2 - R^2) / (2 r D).  alpha has square-root ends at r = |D - R| and r = D + R, and varies on the
// scale |D - R| next to the first: the radial range is cut at those radii and at the kernel knots, refined
// geometrically (ratio 4) away from both square-root points, and each panel takes 16 Gauss-Legendre nodes, in t^2
// (resp. (1 - t)^2) on a panel that ends at a square-root point; the W^5 kernels take panels of at most 1/16.  1 -+ cos(alpha) are formed from the products
// (r - |D - R|)(D + R - r) etc. so that they keep their relative accuracy next to the ends.  A disk that covers the
// support (D <= R - 1) gives the whole-support values, one out of reach (D >= 1 + R) nothing.
//
// Replaces the reference's table lookup of the disk element (edge/warpdisk.py:98-131, its tables built by
// edge/disk.py:113-157 from the divergence theorem on the circle) and the slot fill of scene/fixedadj.py:199-246.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "edge2d_device.h"
#include "wsb_internal.h"

struct wsb_disks2d {
    int device = 0;
    int m = 0;
    double support_max = 0, cell = 0, lo0 = 0, lo1 = 0;
    int nx = 1, ny = 1, K = 1;
    int64_t bytes = 0;
    double4* rec = nullptr;      // [m] (cx, cy, R, 0) by disk index
    int32_t* cstart = nullptr;   // [nx*ny+1]
    int32_t* citems = nullptr;   // [m] disk indices by cell, ascending index within a cell
};

namespace {

constexpr int DISK_NODES = 16;

template <int KID>
__host__ __device__ __forceinline__ void disk_profile(double r, double& P, double& dP) {
    if constexpr (KID == WSB_W2P5 || KID == WSB_W4P5) {
        double radial;
        chord::p5_wr<KID>(r, P, radial);
        dP = -radial * r;
    } else {
        // pi W = u^K q(r), u = HI - r, on the piece of r: P' = u^(K-1) (u q' - K q)
        using T = DTab<KID>;
        const int p = (T::NP == 2 && r >= T::HI[0]) ? 1 : 0;
        double q = T::Q[p][T::DEG], dq = 0.0;
#pragma unroll
        for (int n = T::DEG - 1; n >= 0; --n) {
            dq = dq * r + q;
            q = q * r + T::Q[p][n];
        }
        const int K = T::K[p];
        const double u = T::HI[p] - r;
        double uk1 = 1.0;
        for (int k = 1; k < K; ++k) uk1 *= u;
        const double pv = (K > 0) ? uk1 * u * q : q;
        const double dv = (K > 0) ? uk1 * (u * dq - (double)K * q) : dq;
        P = (r < 1.0) ? pv : 0.0;
        dP = (r < 1.0) ? dv : 0.0;
    }
}

// The Gauss nodes of one disk's radial integrals, panels cut for kernel PKID (its knot, its widest panel), nodes j0,
// j0 + jstep, ... of every panel (the host takes them all, a 16-lane group one node per lane).  Calls inner(r, wg) for
// a node inside the disk where the whole circle counts (alpha = pi) and arc(r, wg, al, sa, sc) on the arc range, with
// al = alpha, sa = sin(alpha), sc = sin(alpha) cos(alpha).  The caller handles D >= 1 + R (nothing) and D <= R - 1
// (the whole support).
template <int PKID, class Inner, class Arc>
__host__ __device__ __forceinline__ void disk_nodes(double D, double R, int j0, int jstep, Inner&& inner, Arc&& arc) {
    using T = DTab<PKID>;
    // the W^5 kernels (degree 25 / 40, every power of u = 1 - r up to 30) take panels of at most 1/16
    constexpr double MAXW = (PKID == WSB_W2P5 || PKID == WSB_W4P5) ? 0.0625 : 2.0;
    const double a = fabs(D - R), b = D + R;
    const bool inside = D < R;
    // r < a inside the disk: the whole circle (alpha = pi); polynomial in r, exact with 16 nodes between knots
    if (inside) {
        const double end = fmin(a, 1.0);
        double lo = 0.0;
        for (int it = 0; it < 64 && lo < end; ++it) {
            double hi = fmin(end, lo + MAXW);
            if (T::NP == 2 && lo < T::HI[0]) hi = fmin(hi, T::HI[0]);
            const double w = hi - lo;
            for (int j = j0; j < DISK_NODES; j += jstep) inner(lo + w * GLTab::X16[j], w * GLTab::W16[j]);
            lo = hi;
        }
    }
    if (a < 1.0) {
        const bool bend = b <= 1.0;                  // the outer square-root point lies inside the support
        const double width = b - a;
        const double send = fmin(b, 1.0) - a;        // the arc range in s = r - a
        double knot = -1.0;                          // kernel knot inside the arc range
        if (T::NP == 2 && T::HI[0] > a && T::HI[0] < fmin(b, 1.0)) knot = T::HI[0] - a;
        const bool tiny = a <= 1e-15 * send;
        double cur = 0.0;
        for (int it = 0; it < 200 && cur < send; ++it) {
            const double nh = (knot > cur) ? knot : send;
            // a panel's other singular points stay at least a third of its width away (half of it for a panel in
            // t^2 or (1 - t)^2, whose map halves the distance near its regular end)
            double nxt;
            int kind;
            if (cur == 0.0) {
                kind = 1;
                nxt = fmin(tiny ? nh : fmin(nh, 2.0 * a), MAXW);
                nxt = (bend && nxt == send) ? 0.5 * send : fmin(nxt, 0.5 * width);
            } else {
                kind = 0;
                nxt = fmin(fmin(nh, 4.0 * cur), cur + MAXW);
                if (bend && nxt == send && send <= 2.0 * cur) kind = 2;
                else nxt = fmin(nxt, 0.25 * (cur + 3.0 * width));
            }
            if (!(nxt > cur) || it == 199) {         // rounding left no room: close the range in one panel
                nxt = send;
                kind = bend ? 2 : 0;
            }
            const double wd = nxt - cur;
            for (int j = j0; j < DISK_NODES; j += jstep) {
                const double t = GLTab::X16[j], wq = GLTab::W16[j];
                double s, bmr, wg;
                if (kind == 1) {
                    s = wd * t * t;
                    bmr = width - s;
                    wg = 2.0 * wd * t * wq;
                } else if (kind == 2) {
                    const double u = 1.0 - t, du = wd * u * u;
                    s = nxt - du;
                    bmr = du;
                    wg = 2.0 * wd * u * wq;
                } else {
                    s = cur + wd * t;
                    bmr = width - s;
                    wg = wd * wq;
                }
                const double r = a + s;
                const double i2rD = 0.5 / (r * D);
                double omc, opc;                    // 1 - cos(alpha), 1 + cos(alpha)
                if (inside) {
                    omc = bmr * (r + a) * i2rD;
                    opc = s * (r + b) * i2rD;
                } else {
                    omc = s * bmr * i2rD;
                    opc = (r + a) * (r + b) * i2rD;
                }
                const double al = 2.0 * atan2(sqrt(omc), sqrt(opc));
                const double sc = sqrt(omc * opc) * 0.5 * (opc - omc);
                const double sa = sqrt(omc * opc);
                arc(r, wg, al, sa, sc);
            }
            cur = nxt;
        }
    }
}

// a node's weights of the seven radial sums below, shared by every kernel: the sums of P take a0, a1, a5, a6, those
// of P' take b2, b3, b4 (a1 = b2 = 0 inside the disk, where the whole circle counts)
struct DiskNode {
    double r, a0, a1, a5, a6, b2, b3, b4;
};

__host__ __device__ __forceinline__ DiskNode disk_inner_node(double r, double wg) {
    const double pi = 3.1415926535897932384626433832795;
    const double g = wg * r * r * pi;
    return DiskNode{r, wg * 2.0 * pi * r, 0.0, g * r, g * r, 0.0, g, g};
}

__host__ __device__ __forceinline__ DiskNode disk_arc_node(double r, double wg, double al, double sa, double sc) {
    const double rr = r * r;
    const double b3 = wg * rr * (al + sc), b4 = wg * rr * (al - sc);
    return DiskNode{r, wg * 2.0 * al * r, wg * 2.0 * sa * rr, b3 * r, b4 * r, wg * 2.0 * sa * r, b3, b4};
}

// one node's terms of kernel KID added to v (ARC: the node lies on the arc range, else inside the disk)
template <int KID, bool ARC>
__host__ __device__ __forceinline__ void disk_node_terms(const DiskNode& q, double v[7]) {
    double P, dP;
    disk_profile<KID>(q.r, P, dP);
    v[0] += P * q.a0;
    if (ARC) {
        v[1] += P * q.a1;
        v[2] -= dP * q.b2;
    }
    v[3] -= dP * q.b3;
    v[4] -= dP * q.b4;
    v[5] += P * q.a5;
    v[6] += P * q.a6;
}

// the whole support (D <= R - 1): the values before the factor 1/pi, on the first node only
template <int KID>
__host__ __device__ __forceinline__ void disk_whole_terms(double v[7]) {
    using T = DTab<KID>;
    v[0] = v[3] = v[4] = T::MASS * 3.1415926535897932384626433832795;
    v[5] = v[6] = T::SECOND * 3.1415926535897932384626433832795;
}

// v: lam, m1 (along e), g0 (along e), g1 (e e), g1 (t t), m2 (e e), m2 (t t) of one disk, units of h, from the
// Gauss nodes j0, j0 + jstep, ... of every panel
template <int KID>
__host__ __device__ void disk_channels_part(double D, double R, int j0, int jstep, double v[7]) {
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = 0.0;
    if (!(D < 1.0 + R)) return;
    if (D <= R - 1.0) {
        if (j0 == 0) disk_whole_terms<KID>(v);
        return;                                      // (the caller's 1/pi gives the whole-support values)
    }
    disk_nodes<KID>(
        D, R, j0, jstep, [&](double r, double wg) { disk_node_terms<KID, false>(disk_inner_node(r, wg), v); },
        [&](double r, double wg, double al, double sa, double sc) {
            disk_node_terms<KID, true>(disk_arc_node(r, wg, al, sa, sc), v);
        });
}

// the scheme kernels of MASK (bit k = kernel id k, cubic .. w6) at the same nodes: the panels of the cubic (cut at
// its knot r = 1/2) when it is in MASK, else the panels that w2, w4 and w6 share; v[k] as disk_channels_part
template <int MASK>
__host__ __device__ void disk_channels_multi_part(double D, double R, int j0, int jstep, double v[4][7]) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < 7; ++c) v[k][c] = 0.0;
    if (!(D < 1.0 + R)) return;
    if (D <= R - 1.0) {
        if (j0 == 0) {
            if (MASK & 1) disk_whole_terms<WSB_CUBIC>(v[0]);
            if (MASK & 2) disk_whole_terms<WSB_W2>(v[1]);
            if (MASK & 4) disk_whole_terms<WSB_W4>(v[2]);
            if (MASK & 8) disk_whole_terms<WSB_W6>(v[3]);
        }
        return;
    }
    constexpr int PKID = (MASK & 1) ? WSB_CUBIC : WSB_W2;
    disk_nodes<PKID>(
        D, R, j0, jstep,
        [&](double r, double wg) {
            const DiskNode q = disk_inner_node(r, wg);
            if (MASK & 1) disk_node_terms<WSB_CUBIC, false>(q, v[0]);
            if (MASK & 2) disk_node_terms<WSB_W2, false>(q, v[1]);
            if (MASK & 4) disk_node_terms<WSB_W4, false>(q, v[2]);
            if (MASK & 8) disk_node_terms<WSB_W6, false>(q, v[3]);
        },
        [&](double r, double wg, double al, double sa, double sc) {
            const DiskNode q = disk_arc_node(r, wg, al, sa, sc);
            if (MASK & 1) disk_node_terms<WSB_CUBIC, true>(q, v[0]);
            if (MASK & 2) disk_node_terms<WSB_W2, true>(q, v[1]);
            if (MASK & 4) disk_node_terms<WSB_W4, true>(q, v[2]);
            if (MASK & 8) disk_node_terms<WSB_W6, true>(q, v[3]);
        });
}

template <int KID>
__host__ __device__ void disk_channels(double D, double R, double v[7]) {
    disk_channels_part<KID>(D, R, 0, 1, v);
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] *= WSB_INV_PI;
}

// one thread per row: the disks of the 3 x 3 cells around the row's cell, cell by cell (x outer, y inner) and in
// cell-list order, kept when |centre - x| < h + R (reference scene/fixedadj.py:199-222)
__global__ void disk_slots_kernel(const double4* __restrict__ rec, const int32_t* __restrict__ cstart,
                                  const int32_t* __restrict__ citems, double lo0, double lo1, double cell, int nx, int ny,
                                  int K, const double* __restrict__ lpos, const double* __restrict__ lsup,
                                  const uint8_t* __restrict__ valid, int n, int32_t* __restrict__ slots,
                                  int32_t* __restrict__ counts) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int32_t* out = slots + (size_t)r * K;
    int cnt = 0;
    if (!valid || valid[r]) {
        const double px = lpos[2 * r], py = lpos[2 * r + 1], h = lsup[r];
        const double fx = floor((px - lo0) / cell), fy = floor((py - lo1) / cell);
        if (fx >= -1.0 && fx <= (double)nx && fy >= -1.0 && fy <= (double)ny) {
            const int cx = (int)fx, cy = (int)fy;
            for (int ddx = -1; ddx <= 1; ++ddx) {
                const int ix = cx + ddx;
                if (ix < 0 || ix >= nx) continue;
                for (int ddy = -1; ddy <= 1; ++ddy) {
                    const int iy = cy + ddy;
                    if (iy < 0 || iy >= ny) continue;
                    const int cid = ix * ny + iy;
                    for (int i = cstart[cid]; i < cstart[cid + 1]; ++i) {
                        const int m = citems[i];
                        const double4 c = rec[m];
                        const double ex = px - c.x, ey = py - c.y;
                        const double ex2 = ex * ex, ey2 = ey * ey;
                        if (sqrt(ex2 + ey2) < h + c.z && cnt < K) out[cnt++] = m;
                    }
                }
            }
        }
    }
    if (counts) counts[r] = cnt;
    for (int s = cnt; s < K; ++s) out[s] = -1;
}

// a group of 16 lanes per row: the row's disks in slot order, each disk's panels with one Gauss node per lane and a
// fixed butterfly sum over the group, the disk's channels turned into the body frame, then the physical scaling, the
// world rotation and the store of finalize_row by the group's first lane (the disk is complete: no indicator term)
// x plus its value in the lane that the DPP control CTRL names within the lane's row of 16 (the two 32-bit halves moved
// by the VALU, no LDS permute)
template <int CTRL>
__device__ __forceinline__ double dpp_pair_sum(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const int ylo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
    const int yhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
    return x + __hiloint2double(yhi, ylo);
}

// the sum over the 16 lanes of a row's group, which is one DPP row (groups are aligned to 16 lanes and all their lanes
// are active together): lane ^ 1, lane ^ 2 (quad permutes), then the mirror within each half-row and within the row.
// Each step adds the same two values in both lanes of a pair, so every lane ends with the same sum.
__device__ __forceinline__ double group_sum(double x) {
    x = dpp_pair_sum<0xB1>(x);      // quad_perm [1, 0, 3, 2]
    x = dpp_pair_sum<0x4E>(x);      // quad_perm [2, 3, 0, 1]
    x = dpp_pair_sum<0x141>(x);     // row_half_mirror: lane i <-> 7 - i of its eight
    x = dpp_pair_sum<0x140>(x);     // row_mirror: lane i <-> 15 - i of its row
    return x;
}

// the twelve channels of one disk from its seven radial sums v (already times 1/pi), centre at (yx, yy) = D e from the
// row, added to c
__device__ __forceinline__ void disk_add_channels(const double v[7], double yx, double yy, double D, double c[wsb::NCH]) {
    double ex = 1.0, ey = 0.0;
    if (D > 1e-12) {
        ex = yx / D;
        ey = yy / D;
    }
    const double tx = -ey, ty = ex;
    c[0] += v[0];
    c[1] += v[1] * ex;
    c[2] += v[1] * ey;
    c[3] += v[2] * ex;
    c[4] += v[2] * ey;
    const double gxy = v[3] * ex * ey + v[4] * tx * ty;
    c[5] += v[3] * ex * ex + v[4] * tx * tx;
    c[6] += gxy;
    c[7] += gxy;
    c[8] += v[3] * ey * ey + v[4] * ty * ty;
    c[9] += v[5] * ex * ex + v[6] * tx * tx;
    c[10] += v[5] * ex * ey + v[6] * tx * ty;
    c[11] += v[5] * ey * ey + v[6] * ty * ty;
}

template <int KID>
__global__ void __launch_bounds__(256) disk_moments_kernel(const double4* __restrict__ rec,
                                                           const int32_t* __restrict__ slots,
                                                           const int32_t* __restrict__ counts, int K,
                                                           const double* __restrict__ lpos,
                                                           const double* __restrict__ lsup, int n,
                                                           const double* __restrict__ rot4, int flags,
                                                           double* __restrict__ out) {
    const int lane = threadIdx.x & (DISK_NODES - 1);
    const int r = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) / DISK_NODES);
    if (r >= n) return;                              // whole groups leave together
    double c[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) c[k] = 0.0;
    const double px = lpos[2 * r], py = lpos[2 * r + 1], h = lsup[r], ih = 1.0 / h;
    const int cnt = counts ? min(counts[r], K) : K;
    for (int s = 0; s < cnt; ++s) {
        const int m = slots[(size_t)r * K + s];
        if (m < 0) continue;
        const double4 d = rec[m];
        const double yx = (d.x - px) * ih, yy = (d.y - py) * ih;
        const double D = sqrt(yx * yx + yy * yy);
        double v[7];
        disk_channels_part<KID>(D, d.z * ih, lane, DISK_NODES, v);
#pragma unroll
        for (int k = 0; k < 7; ++k) v[k] = group_sum(v[k]) * WSB_INV_PI;
        disk_add_channels(v, yx, yy, D, c);
    }
    if (lane == 0) finalize_row(c, 0.0, h, nullptr, rot4, flags | WSB_OUT_NO_INDICATOR, out + (size_t)r * wsb::NCH);
}

struct DiskOuts {
    double* p[4];               // by kernel id WSB_CUBIC .. WSB_W6 (nullptr: not requested)
};

// the scheme kernels of MASK in one pass (wsb_disk_moments2d_multi): per pair every node's geometry (alpha, its sine
// and cosine) once for all of them; after the group sums lane k of the row's group (k < 4) keeps the channel sums of
// kernel k and stores them
template <int MASK>
__global__ void __launch_bounds__(256) disk_moments_multi_kernel(const double4* __restrict__ rec,
                                                                 const int32_t* __restrict__ slots,
                                                                 const int32_t* __restrict__ counts, int K,
                                                                 const double* __restrict__ lpos,
                                                                 const double* __restrict__ lsup, int n,
                                                                 const double* __restrict__ rot4, int flags, DiskOuts outs) {
    const int lane = threadIdx.x & (DISK_NODES - 1);
    const int r = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) / DISK_NODES);
    if (r >= n) return;                              // whole groups leave together
    const int kk = lane & 3;
    double c[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) c[k] = 0.0;
    const double px = lpos[2 * r], py = lpos[2 * r + 1], h = lsup[r], ih = 1.0 / h;
    const int cnt = counts ? min(counts[r], K) : K;
    for (int s = 0; s < cnt; ++s) {
        const int m = slots[(size_t)r * K + s];
        if (m < 0) continue;
        const double4 d = rec[m];
        const double yx = (d.x - px) * ih, yy = (d.y - py) * ih;
        const double D = sqrt(yx * yx + yy * yy);
        double v[4][7];
        disk_channels_multi_part<MASK>(D, d.z * ih, lane, DISK_NODES, v);
        double w[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if ((MASK >> q) & 1) v[q][k] = group_sum(v[q][k]) * WSB_INV_PI;
            w[k] = kk == 0 ? v[0][k] : kk == 1 ? v[1][k] : kk == 2 ? v[2][k] : v[3][k];
        }
        disk_add_channels(w, yx, yy, D, c);
    }
    if (lane < 4 && ((MASK >> lane) & 1))
        finalize_row(c, 0.0, h, nullptr, rot4, flags | WSB_OUT_NO_INDICATOR, outs.p[lane] + (size_t)r * wsb::NCH);
}

// ------------------------------------------------------------------------------------ cone area, direction extreme
// The rays from the query at angle psi from the centre direction meet the disk (centre distance D, radius R, support 1)
// between t-+ = D cos psi -+ q, q = sqrt(R^2 - D^2 sin^2 psi); the area swept between angles 0 and psi is
// I(psi) = int (1/2)(hi^2 - lo^2), hi / lo = t+ / t- clipped to [0, 1].  On a range of psi where neither t+ nor t-
// crosses 0 or 1 the integrand is a constant, 1/2 or (1/2) t-+^2, whose antiderivative is
// (1/2) [R^2 psi + (D^2 / 2) sin 2 psi -+ (s q + R^2 asin(s / R))], s = D sin psi.  The ranges are cut at the tangent
// ray (sin psi = R / D) and at the rays through the intersections of the circle with the support circle.
__device__ inline double disk_sweep_prim(double D, double R, double psi, double sgn) {
    const double s = fmin(fmax(D * sin(psi), -R), R);
    const double q = sqrt(fmax(R * R - s * s, 0.0));
    return 0.5 * (R * R * psi + 0.5 * D * D * sin(2.0 * psi) + sgn * (s * q + R * R * asin(s / R)));
}

// I(psi) for psi in [0, pi]
__device__ inline double disk_sweep(double D, double R, double psi) {
    double cut[4] = {0.0, psi, psi, psi};
    int nc = 1;
    if (D > R) {
        const double a = asin(fmin(R / D, 1.0));
        if (a > 0.0 && a < psi) cut[nc++] = a;
    }
    if (D > 1e-12) {
        const double c = (D * D + 1.0 - R * R) / (2.0 * D);
        if (fabs(c) <= 1.0) {
            const double a = acos(c);
            if (a > 0.0 && a < psi) cut[nc++] = a;
        }
    }
    if (nc == 3 && cut[2] < cut[1]) {
        const double t = cut[1];
        cut[1] = cut[2];
        cut[2] = t;
    }
    cut[nc] = psi;
    double total = 0.0;
    for (int k = 0; k < nc; ++k) {
        const double lo = cut[k], hi = cut[k + 1];
        if (!(hi > lo)) continue;
        const double m = 0.5 * (lo + hi), sm = D * sin(m);
        if (!(fabs(sm) < R)) continue;                 // these rays miss the disk
        const double q = sqrt(R * R - sm * sm), tp = D * cos(m) + q, tm = D * cos(m) - q;
        if (!(tp > 0.0)) continue;                     // the disk lies behind the query on these rays
        // hi part: 1/2 (clipped at the support) or 1/2 t+^2; lo part: 1/2 (disk beyond the support), 1/2 t-^2, or 0
        const double hi_part = tp >= 1.0 ? 0.5 * (hi - lo) : disk_sweep_prim(D, R, hi, 1.0) - disk_sweep_prim(D, R, lo, 1.0);
        double lo_part = 0.0;
        if (tm >= 1.0) lo_part = 0.5 * (hi - lo);
        else if (tm > 0.0) lo_part = disk_sweep_prim(D, R, hi, -1.0) - disk_sweep_prim(D, R, lo, -1.0);
        total += hi_part - lo_part;
    }
    return total;
}

__device__ inline double wrap_angle(double a) {
    const double two_pi = 6.283185307179586476925286766559, pi = 3.1415926535897932384626433832795;
    double r = a + pi;
    r = r - two_pi * floor(r / two_pi);
    return r - pi;
}

// I at any angle: odd, and 2 pi periodic up to the whole area 2 I(pi)
__device__ inline double disk_sweep_any(double D, double R, double psi) {
    const double two_pi = 6.283185307179586476925286766559, pi = 3.1415926535897932384626433832795;
    const double p0 = wrap_angle(psi);
    const double k = floor((psi - p0) / two_pi + 0.5);
    double v = disk_sweep(D, R, fabs(p0));
    if (p0 < 0.0) v = -v;
    if (k != 0.0) v += k * 2.0 * disk_sweep(D, R, pi);
    return v;
}

// one thread per row: area(disks cap disk(p, H) cap wedge(p, axis, half_angle)) added to out[q], the full area to
// out[n + q] (reference scene/fused.py:312-433, the same quantities as the surface form)
__global__ void __launch_bounds__(256) disk_cone_area_kernel(const double4* __restrict__ rec,
                                                             const int32_t* __restrict__ slots,
                                                             const int32_t* __restrict__ counts, int K,
                                                             const double* __restrict__ lpos,
                                                             const double* __restrict__ lsup, int n,
                                                             const double* __restrict__ axes, double pc, double ps,
                                                             double half_angle, double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double pi = 3.1415926535897932384626433832795;
    const double H = lsup[q], px = lpos[2 * q], py = lpos[2 * q + 1];
    const double wx = axes[2 * q], wy = axes[2 * q + 1];
    const double th = atan2(-ps * wx + pc * wy, pc * wx + ps * wy);       // the axis in the body frame
    const bool wedge_on = half_angle < pi;
    double a_wedge = 0.0, a_full = 0.0;
    const int cnt = counts ? min(counts[q], K) : K;
    for (int i = 0; i < cnt; ++i) {
        const int m = slots[(size_t)q * K + i];
        if (m < 0) continue;
        const double4 d = rec[m];
        const double cx = d.x - px, cy = d.y - py;
        const double D = sqrt(cx * cx + cy * cy) / H, R = d.z / H;
        const double full = 2.0 * disk_sweep(D, R, pi);
        a_full += full;
        if (wedge_on) {
            const double g = wrap_angle(th - atan2(cy, cx));
            a_wedge += disk_sweep_any(D, R, g + half_angle) - disk_sweep_any(D, R, g - half_angle);
        } else {
            a_wedge += full;
        }
    }
    out[q] += a_wedge * H * H;
    out[n + q] += a_full * H * H;
}

// one thread per row: the largest |cos| between the world vector and a direction from p to a point of a disk within H,
// max-accumulated into out[q].  The directions to the disk points within H form [-b, b] about the centre direction:
// b = asin(R / d) when the tangent points are within H, else the angle where the circle meets the support circle;
// psi = the angle between the line of the vector and the centre direction: cos(max(0, psi - b)); 1 inside a disk.
__global__ void __launch_bounds__(256) disk_dir_extreme_kernel(const double4* __restrict__ rec,
                                                               const int32_t* __restrict__ slots,
                                                               const int32_t* __restrict__ counts, int K,
                                                               const double* __restrict__ lpos,
                                                               const double* __restrict__ lsup, int n,
                                                               const double* __restrict__ vecs, double pc, double ps,
                                                               double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double H = lsup[q], px = lpos[2 * q], py = lpos[2 * q + 1];
    const double wx = vecs[2 * q], wy = vecs[2 * q + 1];
    const double bx = pc * wx + ps * wy, by = -ps * wx + pc * wy;
    const double nb = sqrt(bx * bx + by * by);
    if (!(nb > 0.0)) return;
    double best = 0.0;
    const int cnt = counts ? min(counts[q], K) : K;
    for (int i = 0; i < cnt; ++i) {
        const int m = slots[(size_t)q * K + i];
        if (m < 0) continue;
        const double4 dk = rec[m];
        const double cx = dk.x - px, cy = dk.y - py, R = dk.z;
        const double d = sqrt(cx * cx + cy * cy);
        double f = 0.0;
        if (d <= R) {
            f = 1.0;
        } else if (d - R < H) {
            const double b = (d * d - R * R <= H * H) ? asin(fmin(R / d, 1.0))
                                                      : acos(fmin(fmax((d * d + H * H - R * R) / (2.0 * d * H), -1.0), 1.0));
            const double psi = acos(fmin(fabs(cx * bx + cy * by) / (d * nb), 1.0));
            f = cos(fmax(0.0, psi - b));
        }
        best = fmax(best, f);
    }
    out[q] = fmax(out[q], best);
}

template <int KID>
void launch_moments(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K, const double* lpos,
                    const double* lsup, int n, const double* rot4, int flags, double* out, hipStream_t st) {
    const size_t threads = (size_t)n * DISK_NODES;
    hipLaunchKernelGGL(disk_moments_kernel<KID>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, g->rec,
                       slots, counts, K, lpos, lsup, n, rot4, flags, out);
}

template <int MASK>
void launch_moments_multi(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K, const double* lpos,
                          const double* lsup, int n, const double* rot4, int flags, const DiskOuts& outs, hipStream_t st) {
    const size_t threads = (size_t)n * DISK_NODES;
    hipLaunchKernelGGL(disk_moments_multi_kernel<MASK>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, g->rec,
                       slots, counts, K, lpos, lsup, n, rot4, flags, outs);
}

template <int MASK>
int dispatch_multi(int mask, const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K, const double* lpos,
                   const double* lsup, int n, const double* rot4, int flags, const DiskOuts& outs, hipStream_t st) {
    if constexpr (MASK > 15) {
        return wsb::set_error(WSB_EINVAL, "no kernel requested");
    } else {
        if (mask == MASK) {
            launch_moments_multi<MASK>(g, slots, counts, K, lpos, lsup, n, rot4, flags, outs, st);
            return WSB_OK;
        }
        return dispatch_multi<MASK + 1>(mask, g, slots, counts, K, lpos, lsup, n, rot4, flags, outs, st);
    }
}

template <int KID>
void host_channels(const double* D, const double* R, int n, double* out) {
    for (int i = 0; i < n; ++i) disk_channels<KID>(D[i], R[i], out + 7 * (size_t)i);
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

// the static cell list of the centres (reference scene/scene.py:869-888): cell = support_max + the largest radius over
// the centres' box grown by one cell, disks by cell (stable, so ascending index within a cell), K = the most disks of
// any 3 x 3 block of cells (at least 1)
extern "C" int wsb_disks2d_create(const double* centres, const double* radii, int m, double support_max,
                                  wsb_disks2d** out, int* K) {
    WSB_REQUIRE(centres && radii && out, WSB_EINVAL, "null argument");
    WSB_REQUIRE(m >= 1, WSB_EINVAL, "a disk array needs at least one disk (m=%d)", m);
    WSB_REQUIRE(support_max > 0 && std::isfinite(support_max), WSB_EINVAL, "support_max must be positive and finite");
    double rmax = 0.0, lo0 = centres[0], lo1 = centres[1], hi0 = centres[0], hi1 = centres[1];
    for (int i = 0; i < m; ++i) {
        WSB_REQUIRE(std::isfinite(centres[2 * i]) && std::isfinite(centres[2 * i + 1]), WSB_EINVAL, "non-finite centre");
        WSB_REQUIRE(radii[i] > 0 && std::isfinite(radii[i]), WSB_EINVAL, "disk %d: radius must be positive and finite", i);
        rmax = std::max(rmax, radii[i]);
        lo0 = std::min(lo0, centres[2 * i]);
        lo1 = std::min(lo1, centres[2 * i + 1]);
        hi0 = std::max(hi0, centres[2 * i]);
        hi1 = std::max(hi1, centres[2 * i + 1]);
    }
    const double cell = support_max + rmax;
    lo0 -= cell;
    lo1 -= cell;
    hi0 += cell;
    hi1 += cell;
    const double dx = std::max(std::floor((hi0 - lo0) / cell) + 1.0, 1.0);
    const double dy = std::max(std::floor((hi1 - lo1) / cell) + 1.0, 1.0);
    WSB_REQUIRE(dx * dy <= (double)(1 << 28), WSB_EINVAL,
                "disk cell grid of %.0f x %.0f cells (support_max %g, largest radius %g) exceeds 2^28 cells", dx, dy,
                support_max, rmax);
    const int nx = (int)dx, ny = (int)dy;
    std::vector<int64_t> cid(m);
    for (int i = 0; i < m; ++i) {
        const int64_t cx = (int64_t)std::floor((centres[2 * i] - lo0) / cell);
        const int64_t cy = (int64_t)std::floor((centres[2 * i + 1] - lo1) / cell);
        cid[i] = cx * ny + cy;
    }
    std::vector<int32_t> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t p, int32_t q) { return cid[p] < cid[q]; });
    const size_t ncell = (size_t)nx * ny;
    std::vector<int32_t> count(ncell, 0), start(ncell + 1, 0);
    for (int i = 0; i < m; ++i) count[cid[i]] += 1;
    for (size_t k = 0; k < ncell; ++k) start[k + 1] = start[k] + count[k];
    int Kmax = 1;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            int tot = 0;
            for (int a = std::max(ix - 1, 0); a <= std::min(ix + 1, nx - 1); ++a)
                for (int b = std::max(iy - 1, 0); b <= std::min(iy + 1, ny - 1); ++b) tot += count[(size_t)a * ny + b];
            Kmax = std::max(Kmax, tot);
        }
    std::vector<double4> rec(m);
    for (int i = 0; i < m; ++i) rec[i] = double4{centres[2 * i], centres[2 * i + 1], radii[i], 0.0};
    auto* g = new wsb_disks2d();
    WSB_HIP(hipGetDevice(&g->device));
    g->m = m;
    g->support_max = support_max;
    g->cell = cell;
    g->lo0 = lo0;
    g->lo1 = lo1;
    g->nx = nx;
    g->ny = ny;
    g->K = Kmax;
    int st = upload(&g->rec, rec.data(), rec.size(), &g->bytes);
    if (!st) st = upload(&g->cstart, start.data(), start.size(), &g->bytes);
    if (!st) st = upload(&g->citems, order.data(), order.size(), &g->bytes);
    if (st) {
        wsb_disks2d_destroy(g);
        return st;
    }
    *out = g;
    if (K) *K = Kmax;
    return WSB_OK;
}

extern "C" int wsb_disks2d_destroy(wsb_disks2d* g) {
    if (!g) return WSB_OK;
    for (void* p : {(void*)g->rec, (void*)g->cstart, (void*)g->citems})
        if (p) (void)hipFree(p);
    delete g;
    return WSB_OK;
}

extern "C" int wsb_disks2d_info(const wsb_disks2d* g, int64_t* info) {
    WSB_REQUIRE(g && info, WSB_EINVAL, "null argument");
    info[0] = g->K;
    info[1] = g->nx;
    info[2] = g->ny;
    info[3] = g->m;
    info[4] = g->bytes;
    return WSB_OK;
}

extern "C" int wsb_disk_adjacency2d(const wsb_disks2d* g, const double* lpos, const double* lsup, const uint8_t* valid,
                                    int n, int32_t* slots, int32_t* counts, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null disk array");
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative row count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(lpos && lsup && slots, WSB_EINVAL, "null buffer");
    hipLaunchKernelGGL(disk_slots_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), g->rec,
                       g->cstart, g->citems, g->lo0, g->lo1, g->cell, g->nx, g->ny, g->K, lpos, lsup, valid, n, slots,
                       counts);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_disk_moments2d(const wsb_disks2d* g, int kernel_id, const int32_t* slots, const int32_t* counts,
                                  int K, const double* lpos, const double* lsup, int n, const double* rot4, int flags,
                                  double* out, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null disk array");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes (n=%d K=%d)", n, K);
    WSB_REQUIRE((size_t)n * DISK_NODES / 256 < (size_t)1 << 31, WSB_EINVAL, "too many rows (%d)", n);
    WSB_REQUIRE(!(flags & ~(WSB_OUT_ACCUMULATE | WSB_OUT_PHYSICAL | WSB_OUT_WORLD | WSB_OUT_NO_INDICATOR)), WSB_EINVAL,
                "unknown flags %d", flags);
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD) || rot4, WSB_EINVAL, "WSB_OUT_WORLD needs rot4");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && lpos && lsup && out, WSB_EINVAL, "null buffer");
    const hipStream_t st = wsb::as_stream(stream);
    switch (kernel_id) {
        case WSB_CUBIC: launch_moments<WSB_CUBIC>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_W2: launch_moments<WSB_W2>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_W4: launch_moments<WSB_W4>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_W6: launch_moments<WSB_W6>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_CONE: launch_moments<WSB_CONE>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_LW2: launch_moments<WSB_LW2>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_LW4: launch_moments<WSB_LW4>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_W2P5: launch_moments<WSB_W2P5>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        case WSB_W4P5: launch_moments<WSB_W4P5>(g, slots, counts, K, lpos, lsup, n, rot4, flags, out, st); break;
        default: return wsb::set_error(WSB_EINVAL, "unknown kernel id %d", kernel_id);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_disk_moments2d_multi(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K,
                                        const double* lpos, const double* lsup, int n, const double* rot4, int flags,
                                        double* const* outs, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null disk array");
    WSB_REQUIRE(outs, WSB_EINVAL, "null output table");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes (n=%d K=%d)", n, K);
    WSB_REQUIRE((size_t)n * DISK_NODES / 256 < (size_t)1 << 31, WSB_EINVAL, "too many rows (%d)", n);
    WSB_REQUIRE(!(flags & ~(WSB_OUT_ACCUMULATE | WSB_OUT_PHYSICAL | WSB_OUT_WORLD | WSB_OUT_NO_INDICATOR)), WSB_EINVAL,
                "unknown flags %d", flags);
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD) || rot4, WSB_EINVAL, "WSB_OUT_WORLD needs rot4");
    DiskOuts o{};
    int mask = 0;
    for (int k = 0; k < 4; ++k) {
        o.p[k] = outs[k];
        if (outs[k]) mask |= 1 << k;
    }
    WSB_REQUIRE(mask, WSB_EINVAL, "no kernel requested");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && lpos && lsup, WSB_EINVAL, "null buffer");
    const int st = dispatch_multi<1>(mask, g, slots, counts, K, lpos, lsup, n, rot4, flags, o, wsb::as_stream(stream));
    if (st) return st;
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_disk_channels_host(int kernel_id, const double* D, const double* R, int n, double* out) {
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(D && R && out, WSB_EINVAL, "null buffer");
    switch (kernel_id) {
        case WSB_CUBIC: host_channels<WSB_CUBIC>(D, R, n, out); break;
        case WSB_W2: host_channels<WSB_W2>(D, R, n, out); break;
        case WSB_W4: host_channels<WSB_W4>(D, R, n, out); break;
        case WSB_W6: host_channels<WSB_W6>(D, R, n, out); break;
        case WSB_CONE: host_channels<WSB_CONE>(D, R, n, out); break;
        case WSB_LW2: host_channels<WSB_LW2>(D, R, n, out); break;
        case WSB_LW4: host_channels<WSB_LW4>(D, R, n, out); break;
        case WSB_W2P5: host_channels<WSB_W2P5>(D, R, n, out); break;
        case WSB_W4P5: host_channels<WSB_W4P5>(D, R, n, out); break;
        default: return wsb::set_error(WSB_EINVAL, "unknown kernel id %d", kernel_id);
    }
    return WSB_OK;
}

extern "C" int wsb_disk_cone_area2d(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K,
                                    const double* lpos, const double* lsup, int n, const double* axes, double pose_cos,
                                    double pose_sin, double half_angle, double* out, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null disk array");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes (n=%d K=%d)", n, K);
    WSB_REQUIRE(half_angle >= 0.0, WSB_EINVAL, "negative half angle");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && lpos && lsup && axes && out, WSB_EINVAL, "null buffer");
    hipLaunchKernelGGL(disk_cone_area_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), g->rec,
                       slots, counts, K, lpos, lsup, n, axes, pose_cos, pose_sin, half_angle, out);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_disk_dir_extreme2d(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K,
                                      const double* lpos, const double* lsup, int n, const double* vecs, double pose_cos,
                                      double pose_sin, double* out, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null disk array");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes (n=%d K=%d)", n, K);
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && lpos && lsup && vecs && out, WSB_EINVAL, "null buffer");
    hipLaunchKernelGGL(disk_dir_extreme_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), g->rec,
                       slots, counts, K, lpos, lsup, n, vecs, pose_cos, pose_sin, out);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}
