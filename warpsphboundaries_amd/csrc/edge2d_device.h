// Device arithmetic of the 2-D edge reductions: the plan interpreter of the plain FP64 form and the
// closed form with shared primitive tables of the optimized forms.
#pragma once

#include <hip/hip_runtime.h>

#include "tables.inc"
#include "wsb_internal.h"

#define WSB_INV_PI 0.31830988618379067154

// ================================================================================================= plain form
// A term-by-term interpreter of the flattened plan, one (query, edge) pair per thread: every term recomputes
// its primitives from scratch (reference edge/warp2d.py:28-149 and edge/warpbc.py:578-638).
namespace plain {

__device__ inline double ipow(double x, int k) {
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

__device__ inline double asinh_ref(double u) {
    const double a = fabs(u);
    const double v = log(a + sqrt(a * a + 1.0));
    return u < 0.0 ? -v : v;
}

// odd antiderivative of r^m, m >= 0 (parity-chain recurrence, z = 0 guarded)
__device__ inline double Im(int m, double s, double z) {
    const double zz = z * z;
    const double r2 = s * s + zz;
    const double r = r2 > 0.0 ? sqrt(r2) : 0.0;
    double val = s;
    int idx = 0;
    if (m % 2 == 1) {
        const double tail = z != 0.0 ? zz * asinh_ref(s / fabs(z)) : 0.0;
        val = (s * r + tail) / 2.0;
        idx = 1;
    }
    while (idx < m) {
        idx += 2;
        const double rp = (idx % 2 == 0) ? ipow(r2, idx / 2) : r * ipow(r2, (idx - 1) / 2);
        val = (s * rp + (double)idx * zz * val) / (double)(idx + 1);
    }
    return val;
}

__device__ inline double binom(int n, int k) {
    double v = 1.0;
    for (int i = 0; i < k; ++i) v = v * (double)(n - i) / (double)(i + 1);
    return v;
}

// int s^j r^m ds, m >= 0
__device__ inline double S(int j, int m, double s, double z) {
    const double zz = z * z;
    double tot = 0.0;
    if (j % 2 == 0) {
        const int h = j / 2;
        for (int i = 0; i <= h; ++i) {
            const double sg = ((h - i) % 2 == 1) ? -1.0 : 1.0;
            tot += binom(h, i) * sg * ipow(zz, h - i) * Im(m + 2 * i, s, z);
        }
    } else {
        const int h = (j - 1) / 2;
        const double r2 = s * s + zz;
        const double r = r2 > 0.0 ? sqrt(r2) : 0.0;
        for (int i = 0; i <= h; ++i) {
            const int mm = m + 2 * i;
            const double rp = (mm % 2 == 0) ? ipow(r2, (mm + 2) / 2) : r * ipow(r2, (mm + 1) / 2);
            const double sg = ((h - i) % 2 == 1) ? -1.0 : 1.0;
            tot += binom(h, i) * sg * ipow(zz, h - i) * rp / (double)(mm + 2);
        }
    }
    return tot;
}

__device__ inline double sdiff(int j, int m, double lo, double hi, double z) { return S(j, m, hi, z) - S(j, m, lo, z); }

__device__ inline double dangle(double z, double lo, double hi) {
    if (z == 0.0) return 0.0;
    const double y = z * (hi - lo);
    const double x = z * z + lo * hi;
    if (y == 0.0 && x == 0.0) return 0.0;
    return atan2(y, x);
}

// coefficient of s^j in (z n0 + s t0)^a (z n1 + s t1)^b
__device__ inline double ypoly_coef(int a, int b, int j, double z, double n0, double n1, double t0, double t1) {
    double tot = 0.0;
    for (int ja = 0; ja <= a; ++ja) {
        const int jb = j - ja;
        if (jb >= 0 && jb <= b) {
            const double ca = binom(a, ja) * ipow(t0, ja) * ipow(z * n0, a - ja);
            const double cb = binom(b, jb) * ipow(t1, jb) * ipow(z * n1, b - jb);
            tot += ca * cb;
        }
    }
    return tot;
}

__device__ inline double edge_integral(int a, int b, double lo, double hi, double z, double n0, double n1, double t0,
                                       double t1, const int32_t* cn, const double* cc, int c0, int c1) {
    double tot = 0.0;
    for (int j = 0; j <= a + b; ++j) {
        const double yc = ypoly_coef(a, b, j, z, n0, n1, t0, t1);
        if (yc != 0.0) {
            double inner = 0.0;
            for (int k = c0; k < c1; ++k) inner += cc[k] * sdiff(j, cn[k], lo, hi, z);
            tot += yc * inner;
        }
    }
    return tot;
}

struct Plan {
    const double* radii;
    const int32_t* cn;
    const double* cc;
    const int32_t* E;   // [nE*7] ch, i, a, b, R, c0, c1
    const int32_t* V;   // [nV*4] ch, R, c0, c1
    const double* mR;
    int nE, nV;
};

// the edge-local channels of one pair (units of h, x 1/pi), p and q relative to the query in units of h
__device__ inline void pair_channels(const Plan& pl, double px, double py, double qx, double qy, double ch[wsb::NCH]) {
    const double dx = qx - px, dy = qy - py;
    const double ell = sqrt(dx * dx + dy * dy);
    if (ell == 0.0) return;
    const double t0 = dx / ell, t1 = dy / ell;
    const double n0 = t1, n1 = -t0;
    const double z = wsb_cross2(px, py, qx, qy) / ell;
    const double s0 = (dx * px + dy * py) / ell;
    const double s1 = (dx * qx + dy * qy) / ell;
    const double az = fabs(z);
    for (int t = 0; t < pl.nE; ++t) {
        const int32_t* e = pl.E + 7 * t;
        const double R = pl.radii[e[4]];
        if (az < R) {
            const double L = sqrt((R - az) * (R + az));
            const double lo = fmax(s0, -L), hi = fmin(s1, L);
            if (lo < hi) {
                const double ni = e[1] == 1 ? n1 : n0;
                ch[e[0]] += ni * edge_integral(e[2], e[3], lo, hi, z, n0, n1, t0, t1, pl.cn, pl.cc, e[5], e[6]);
            }
        }
    }
    for (int t = 0; t < pl.nV; ++t) {
        const int32_t* v = pl.V + 4 * t;
        const double R = pl.radii[v[1]];
        if (az < R) {
            const double L = sqrt((R - az) * (R + az));
            const double lo = fmax(s0, -L), hi = fmin(s1, L);
            if (lo < hi) {
                double poly = 0.0;
                for (int k = v[2]; k < v[3]; ++k) poly += pl.cc[k] / (double)(pl.cn[k] + 2) * sdiff(0, pl.cn[k], lo, hi, z);
                ch[v[0]] += z * poly - pl.mR[t] * dangle(z, lo, hi);
            }
        }
    }
    for (int k = 0; k < wsb::NCH; ++k) ch[k] *= WSB_INV_PI;
}

}  // namespace plain

// ================================================================================================= closed form
// One pair, all twelve channels from ONE set of primitives per kernel block: the chord, the endpoint radii,
// one logarithm (the asinh difference), one angle, and the differences dI_m = I_m(hi) - I_m(lo) of the odd
// antiderivatives of r^m by the linear recurrence  dI_m = (d[s r^m] + m z^2 dI_{m-2}) / (m + 1).  With
//     A[f] = int f(r) ds,  V[f] = int sum f_n r^n/(n+2) ds,  B[f] = int s f(r) ds = sum f_n d[r^(n+2)]/(n+2)
// for the block profile P and its compact potential Phi (exact tables of tables.inc) the channels are
//     lam += z V[P] - mR[P] theta                   m10 += n_x A[Phi]       m01 += n_y A[Phi]
//     G_j -= n_j A[P]                               g1_dj += -n_j (z n_d A[P] + t_d B[P]) + delta_dj lam
//     m20 += n_x (z n_x A[Phi] + t_x B[Phi]) - (z V[Phi] - mR[Phi] theta)
//     m11 += n_x (z n_y A[Phi] + t_y B[Phi])        m02 += n_y (z n_y A[Phi] + t_y B[Phi]) - (z V[Phi] - mR[Phi] theta)
// which is the compact-potential recursion of the reference (edge/np2d.py:46-101, edge/warpbc.py:119-136) with
// the line integrals of every channel sharing their primitives.
namespace closed {

// the two arithmetic forms
enum { FP64 = 1, MIXED = 2 };

__device__ __forceinline__ double dangle_fp64(double z, double lo, double hi) {
    if (z == 0.0) return 0.0;
    const double y = z * (hi - lo);
    const double x = z * z + lo * hi;
    if (y == 0.0 && x == 0.0) return 0.0;
    return atan2(y, x);
}

// the angle through the FP32 hardware path: an exact power-of-two rescale keeps both arguments in the
// normal FP32 range; the angle enters with an O(1) weight, so its FP32 rounding is not amplified
__device__ __forceinline__ double dangle_mixed(double z, double lo, double hi) {
    if (z == 0.0) return 0.0;
    double y = z * (hi - lo);
    double x = z * z + lo * hi;
    if (y == 0.0 && x == 0.0) return 0.0;
    int e;
    (void)frexp(fmax(fabs(x), fabs(y)), &e);
    y = ldexp(y, -e);
    x = ldexp(x, -e);
    return (double)atan2f((float)y, (float)x);
}

// the same two functions without a branch (the fused multi-kernel pass, where a branch splits the slot's code and keeps
// the scheduler from interleaving the chord's transcendental chains with the kernels' sums): the angle's special
// cases by a select; the square root of a >= 0 from an FP32 reciprocal-square-root seed of a scaled by an even power
// of two into [0.5, 2) (exact), one FP64 Newton step, scaled back (relative error ~1e-14 for every finite a > 0; 0
// at a = 0)
__device__ __forceinline__ double dangle_mixed_nb(double z, double lo, double hi) {
    const double y = z * (hi - lo);
    const double x = z * z + lo * hi;
    int e;
    (void)frexp(fmax(fabs(x), fabs(y)), &e);
    const double t = (double)atan2f((float)ldexp(y, -e), (float)ldexp(x, -e));
    return (z == 0.0 || (y == 0.0 && x == 0.0)) ? 0.0 : t;
}

__device__ __forceinline__ double xsqrt_mixed_nb(double a) {
    int e;
    (void)frexp(a, &e);
    const int k = e >> 1;
    const double s = ldexp(a, -2 * k);
    const double g = (double)__builtin_amdgcn_rsqf((float)s);
    const double y = s * g;
    const double r = ldexp(fma(fma(-y, y, s), 0.5 * g, y), k);
    return a > 0.0 ? r : 0.0;
}

template <int FORM>
__device__ __forceinline__ double xsqrt(double a) {
    if (FORM == MIXED) {
        // FP32 reciprocal square root seed, one FP64 Newton step (relative error ~1e-14)
        if (a > 1e-30 && a < 1e30) {
            const double g = (double)__builtin_amdgcn_rsqf((float)a);
            const double y = a * g;
            return fma(fma(-y, y, a), 0.5 * g, y);
        }
    }
    return sqrt(a);
}

// One (query, edge) pair, all twelve channels.  Inputs: the rounded differences u = a - x, v = b - x of the edge
// endpoints and the query (the numerators of the reference's relative coordinates), the support h and 1/h, and
// the edge's unit tangent t and 1/length in the local frame.
//
// The sign of z (and whether it is zero) must be the sign of the reference's rounded predicate
// cross2((a-x)/h, (b-x)/h) that the indicator uses: the cross product of u and v decides it when it exceeds
// 16 u_r times the magnitude of its products (u_r = 2^-53; the divisions by h and the roundings move the
// predicate by at most 5 u_r of that magnitude), else the reference's predicate itself is evaluated.
template <int KID, int FORM>
__device__ __forceinline__ void pair_channels(double ux, double uy, double vx, double vy, double h, double ih, double tx,
                                              double ty, double iell, double c[wsb::NCH]) {
    using T = KTab<KID>;
    constexpr int NM = T::NMAXP;
    if (iell == 0.0) return;
    double cr = wsb_cross2(ux, uy, vx, vy);
    const double mag = fabs(ux * vy) + fabs(uy * vx);
    if (!(mag > 1e-280 && fabs(cr) > 1.7763568394002505e-15 * mag)) cr = wsb_cross2(ux / h, uy / h, vx / h, vy / h) * h * h;
    const double z = cr * ih * iell;                     // cross(p', q') / |q' - p'|
    const double px = ux * ih, py = uy * ih, qx = vx * ih, qy = vy * ih;
    const double s0 = tx * px + ty * py, s1 = tx * qx + ty * qy;
    const double n0 = ty, n1 = -tx;
    const double az = fabs(z), zz = z * z;
    const double r0sq = px * px + py * py, r1sq = qx * qx + qy * qy;
    double r0 = -1.0, r1 = -1.0;                         // endpoint radii, computed once when an unclipped end needs them
#pragma unroll
    for (int b = 0; b < T::NB; ++b) {
        const double R = T::R[b];
        if (az < R) {
            const bool clo = s0 < 0.0 && r0sq > R * R;
            const bool chi = s1 > 0.0 && r1sq > R * R;
            const double L = (clo || chi) ? xsqrt<FORM>((R - az) * (R + az)) : 0.0;
            const double lo = clo ? -L : s0, hi = chi ? L : s1;
            if (lo < hi) {
                if (!clo && r0 < 0.0) r0 = xsqrt<FORM>(r0sq);
                if (!chi && r1 < 0.0) r1 = xsqrt<FORM>(r1sq);
                const double rl = clo ? R : r0, rh = chi ? R : r1;
                double dA = 0.0;
                if (z != 0.0) {
                    // asinh(hi/|z|) - asinh(lo/|z|) = log(F(hi)/F(lo)),  F(s) = s + r  (= z^2/(r - s) for s < 0)
                    const double Nh = hi >= 0.0 ? hi + rh : zz, Dh = hi >= 0.0 ? 1.0 : rh - hi;
                    const double Nl = lo >= 0.0 ? lo + rl : zz, Dl = lo >= 0.0 ? 1.0 : rl - lo;
                    dA = log((Nh * Dl) / (Dh * Nl));
                }
                double dI[NM + 1];
                dI[0] = hi - lo;
                if (NM >= 1) dI[1] = 0.5 * (hi * rh - lo * rl + zz * dA);
                double AP = 0.0, VP = 0.0, AF = 0.0, VF = 0.0, BP = 0.0, BF = 0.0;
                double ph = rh, pl = rl;
#pragma unroll
                for (int m = 0; m <= NM + 2; ++m) {
                    if (m >= 2) {
                        ph *= rh;
                        pl *= rl;
                    }
                    if (m >= 2 && m <= NM) dI[m] = ((hi * ph - lo * pl) + (double)m * zz * dI[m - 2]) * (1.0 / (double)(m + 1));
                    if (m <= NM) {
                        if (T::P[b][m] != 0.0) AP += T::P[b][m] * dI[m];
                        if (T::QP[b][m] != 0.0) VP += T::QP[b][m] * dI[m];
                        if (T::F[b][m] != 0.0) AF += T::F[b][m] * dI[m];
                        if (T::QF[b][m] != 0.0) VF += T::QF[b][m] * dI[m];
                    }
                    if (m >= 2) {
                        const double dR = ph - pl;
                        if (T::QP[b][m - 2] != 0.0) BP += T::QP[b][m - 2] * dR;
                        if (T::QF[b][m - 2] != 0.0) BF += T::QF[b][m - 2] * dR;
                    }
                }
                const double th = FORM == MIXED ? dangle_mixed(z, lo, hi) : dangle_fp64(z, lo, hi);
                const double lam = z * VP - T::MRP[b] * th;
                const double valF = z * VF - T::MRF[b] * th;
                const double zn0 = z * n0, zn1 = z * n1;
                const double YxP = zn0 * AP + tx * BP, YyP = zn1 * AP + ty * BP;
                const double YxF = zn0 * AF + tx * BF, YyF = zn1 * AF + ty * BF;
                c[0] += lam;
                c[1] += n0 * AF;
                c[2] += n1 * AF;
                c[3] -= n0 * AP;
                c[4] -= n1 * AP;
                c[5] += lam - n0 * YxP;
                c[6] -= n1 * YxP;
                c[7] -= n0 * YyP;
                c[8] += lam - n1 * YyP;
                c[9] += n0 * YxF - valF;
                c[10] += n0 * YyF;
                c[11] += n1 * YyF - valF;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) c[k] *= WSB_INV_PI;
}

}  // namespace closed

// ================================================================================================= W^5 kernels
// The tensile kernels (W^5 of w2 / w4, degree 25 / 40) are used for their gradient channels only (reference
// scene/tensile.py: T = (1/5) grad int_solid W^5).  In monomials their coefficients reach 1e10 and the closed form
// cancels to 1e-3 (which is why the reference gives them a Chebyshev plan); here the gradient's edge term
// -n int_chord W ds is integrated directly: W >= 0 evaluated as a power of the positive Wendland shape, on 24-node
// Gauss-Legendre panels that double in width away from the foot of the perpendicular (each panel stays at a fixed
// ratio from the branch points s = +-i z of r = sqrt(z^2 + s^2); through the foot, z = 0, W is a polynomial of
// degree <= 40 in s and one panel is exact).
namespace chord {

// pi W5(r) and -(pi W5)'(r) / r from the Wendland shape s (w2: u^4 (1 + 4r), w4: u^6 (1 + 6r + 35/3 r^2), u = 1 - r),
// W5 = c5 s^5: every factor positive, so no cancellation (the 40 powers of u alone cancel by 1e9 near r = 0 for w4);
// -W5'/r = 5 c5 s^4 (-s'/r) with -s'/r = 20 u^3 (w2), 56/3 (1 + 5r) u^5 (w4)
template <int KID>
__host__ __device__ inline void p5_wr(double r, double& W, double& radial) {
    const double u = 1.0 - r;
    if (!(u > 0.0)) {
        W = radial = 0.0;
        return;
    }
    double s, ds;
    if constexpr (KID == WSB_W2P5) {
        const double u3 = u * u * u;
        s = u3 * u * (1.0 + 4.0 * r);
        ds = 20.0 * u3;
    } else {
        const double u2 = u * u, u5 = u2 * u2 * u;
        s = u5 * u * (1.0 + 6.0 * r + (35.0 / 3.0) * r * r);
        ds = (56.0 / 3.0) * (1.0 + 5.0 * r) * u5;
    }
    const double s2 = s * s, s4 = s2 * s2;
    W = P5Tab<KID>::C * s4 * s;
    radial = P5Tab<KID>::C * 5.0 * s4 * ds;
}

template <int KID>
__device__ __forceinline__ double wpi(double r) {
    double W, radial;
    p5_wr<KID>(r, W, radial);
    return W;
}

// int_a^b pi W(sqrt(z^2 + s^2)) ds for 0 <= a < b (units of h)
template <int KID>
__device__ inline double half_chord(double az, double a, double b) {
    if (az < 1e-17 * b) az = 0.0;            // then r = s to within 1e-34 relative on the whole interval
    double sum = 0.0, p = a;
    for (int it = 0; it < 80 && p < b; ++it) {
        const double w = az > 0.0 ? fmax(p, az) : b - p;
        const double q = fmin(b, p + w);
        double part = 0.0;
        for (int j = 0; j < 24; ++j) {
            const double s = p + (q - p) * GLTab::X24[j];
            part += GLTab::W24[j] * wpi<KID>(sqrt(az * az + s * s));
        }
        sum += (q - p) * part;
        p = q;
    }
    return sum;
}

// the gradient channels (3, 4) of one (query, edge) pair; the arguments as closed::pair_channels
template <int KID>
__device__ __forceinline__ void pair_grad(double ux, double uy, double vx, double vy, double h, double ih, double tx,
                                          double ty, double iell, double c[wsb::NCH]) {
    if (iell == 0.0) return;
    const double z = (ux * vy - uy * vx) * ih * iell;
    const double az = fabs(z);
    if (!(az < 1.0)) return;
    const double s0 = (tx * ux + ty * uy) * ih, s1 = (tx * vx + ty * vy) * ih;
    const double L = sqrt((1.0 - az) * (1.0 + az));
    const double lo = fmax(s0, -L), hi = fmin(s1, L);
    if (!(lo < hi)) return;
    double A;
    if (lo >= 0.0) A = half_chord<KID>(az, lo, hi);
    else if (hi <= 0.0) A = half_chord<KID>(az, -hi, -lo);
    else A = half_chord<KID>(az, 0.0, -lo) + half_chord<KID>(az, 0.0, hi);
    c[3] -= ty * A * WSB_INV_PI;
    c[4] += tx * A * WSB_INV_PI;
}

}  // namespace chord

// ================================================================================================= row output
template <int KID>
__device__ __forceinline__ const double* plane_of() { return TTab<KID>::PLANE; }

// indicator x plane vector, physical scaling, world rotation, store / accumulate one row of twelve channels
__device__ __forceinline__ void finalize_row(double c[wsb::NCH], double ind, double h, const double* plane, const double* rot4,
                                             int flags, double* out) {
    if (!(flags & WSB_OUT_NO_INDICATOR)) {
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) c[k] += ind * plane[k];
    }
    if (flags & WSB_OUT_PHYSICAL) {
        const double hh = h * h, ih = 1.0 / h;
        c[1] *= h;
        c[2] *= h;
        c[3] *= ih;
        c[4] *= ih;
        c[9] *= hh;
        c[10] *= hh;
        c[11] *= hh;
    }
    if ((flags & WSB_OUT_WORLD) && rot4) {
        const double r00 = rot4[0], r01 = rot4[1], r10 = rot4[2], r11 = rot4[3];
        // vectors v' = R v
        double a0 = c[1], a1 = c[2];
        c[1] = r00 * a0 + r01 * a1;
        c[2] = r10 * a0 + r11 * a1;
        a0 = c[3];
        a1 = c[4];
        c[3] = r00 * a0 + r01 * a1;
        c[4] = r10 * a0 + r11 * a1;
        // tensors M' = R M R^T, M = [[c5, c6], [c7, c8]]
        const double m00 = c[5], m01 = c[6], m10 = c[7], m11 = c[8];
        c[5] = r00 * (m00 * r00 + m01 * r01) + r01 * (m10 * r00 + m11 * r01);
        c[6] = r00 * (m00 * r10 + m01 * r11) + r01 * (m10 * r10 + m11 * r11);
        c[7] = r10 * (m00 * r00 + m01 * r01) + r11 * (m10 * r00 + m11 * r01);
        c[8] = r10 * (m00 * r10 + m01 * r11) + r11 * (m10 * r10 + m11 * r11);
        const double xx = c[9], xy = c[10], yy = c[11];
        c[9] = r00 * r00 * xx + 2.0 * r00 * r01 * xy + r01 * r01 * yy;
        c[10] = r00 * r10 * xx + (r00 * r11 + r01 * r10) * xy + r01 * r11 * yy;
        c[11] = r10 * r10 * xx + 2.0 * r10 * r11 * xy + r11 * r11 * yy;
    }
    if (flags & WSB_OUT_ACCUMULATE) {
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) out[k] += c[k];
    } else {
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) out[k] = c[k];
    }
}
