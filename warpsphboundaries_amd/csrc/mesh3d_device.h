// Arithmetic of the 3-D face closed form (closed oriented triangle surfaces), shared by the device kernels and
// the host evaluation entry points (the same template code compiled for both sides).
//
// Per (query, face) pair and kernel block (radius R, radial functions P, Phi = -int_r^R t P, Phi2 = -int_r^R t Phi,
// psi_g = sum g_m r^m/(m+3)) the face integral of a radial function g restricted to the support ball is
//
//     S_f[g] = sum_e w_e [ sum_m g_m/(m+2) dK_(m+2)  -  G(R) dK_0 ]  +  (G(R) - G(a)) Theta_f,
//     G(x) = sum_m g_m x^(m+2)/(m+2),
//
// over the in-ball chord of every edge e (w_e = in-plane distance of the edge line from the foot point, a = |z|),
// with K_m = int r^m/(s^2+w^2) ds built by K_m = I_(m-2) + a^2 K_(m-2) from the chord integrals I_k = int r^k ds,
// K_0 = atan(s/|w|)/|w| and a K_-1 = atan(a s/(|w| r))/|w|; Theta_f is the angle of the triangle at the foot point
// (2 pi inside, 0 outside).  The edge line integrals of Phi and Phi2 reuse the same I_k.  This is the edge-cut
// decomposition of the reference's face scalar (mesh3d/reference.py:127-150) with the K primitives of :74-88,
// regrouped so that every channel shares one set of primitives per (pair, block, edge); the channels follow the
// face/volume recursion of mesh3d/reference.py:173-262 summed over the profile.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "tables.inc"

namespace m3 {

constexpr double PI = 3.14159265358979323846;
constexpr double TWO_PI = 6.28318530717958647692;
constexpr double INV_PI = 0.31830988618379067154;
constexpr int NACC = 23;     // M0, M1[3], V_Phi, M2[6], G[3], g1[9] (face sums before the winding term)
constexpr int NOUT = 22;     // baseline(3)[:22]
constexpr int FREC = 36;     // doubles per face record

enum { FP64 = 1, MIXED = 2 };

// face record: v0 v1 v2 (9), unit normal (3), unit edge tangents t_e (9), in-plane outward edge normals mu_e = t_e x n
// (9), edge lengths (3), |(v1 - v0) x (v2 - v0)| (1), sliver flag (1: twice the area at most 1e-2 of the longest edge
// squared), padding (1).  Edge e runs from vertex e to vertex (e+1) % 3.
enum { FR_V = 0, FR_N = 9, FR_T = 12, FR_MU = 21, FR_LEN = 30, FR_NLEN = 33, FR_SLIVER = 34 };

// Angles stay FP64 in both forms.  In the face closed form every angle is weighted by a partial sum of the kernel
// polynomial (its odd part through 1/(r + a), its even part through the clipped pieces), and those partial sums
// exceed the face integral by the growth of the polynomial at -a ((1 + a)^8 for w6): an FP32 angle (1e-7) turns
// into a 1e-4..1e-3 error of the moments there, measured on a 20,480-face sphere at h = 2 mean edge lengths.  The
// mixed form therefore differs from the FP64 form only by FP32-seeded square roots and the cheaper solid-angle
// formula of omega_face.
template <int FORM>
__host__ __device__ inline double angle2(double y, double x) {
    if (y == 0.0 && x == 0.0) return 0.0;
    return atan2(y, x);
}

template <int FORM>
__host__ __device__ inline double xsqrt(double v) {
    if (FORM == MIXED && v > 1e-30 && v < 1e30) {
        // FP32 reciprocal square-root seed, one FP64 Newton step (relative error ~1e-14)
#if defined(__HIP_DEVICE_COMPILE__)
        const double g = (double)__builtin_amdgcn_rsqf((float)v);
#else
        const double g = (double)(1.0f / sqrtf((float)v));
#endif
        const double y = v * g;
        return fma(fma(-y, y, v), 0.5 * g, y);
    }
    return sqrt(v);
}

__host__ __device__ inline double sgn(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

// the face id stored in the bits of a float record field
__host__ __device__ inline int bits_int(float v) {
    int i;
    __builtin_memcpy(&i, &v, sizeof(int));
    return i;
}

// double-double arithmetic (an unevaluated sum hi + lo): exact differences and near-exact triple products for the
// pairs whose plane offset and edge distances are both tiny (the query near an edge of the face).  Contraction is
// off inside these helpers: a product fused into the following sum would count the product's rounding error twice.
struct D2 {
    double hi, lo;
};
__host__ __device__ inline D2 two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ inline D2 quick_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    return {s, b - (s - a)};
}
__host__ __device__ inline D2 dd_add(D2 x, D2 y) {
#pragma clang fp contract(off)
    const D2 s = two_sum(x.hi, y.hi);
    return quick_sum(s.hi, s.lo + (x.lo + y.lo));
}
__host__ __device__ inline D2 dd_sub(D2 x, D2 y) { return dd_add(x, {-y.hi, -y.lo}); }
__host__ __device__ inline D2 dd_mul(D2 x, D2 y) {
#pragma clang fp contract(off)
    const double p = x.hi * y.hi;
    const double e = fma(x.hi, y.hi, -p);
    return quick_sum(p, e + (x.hi * y.lo + x.lo * y.hi));
}
__host__ __device__ inline void dd_cross(const D2 a[3], const D2 b[3], D2 c[3]) {
    c[0] = dd_sub(dd_mul(a[1], b[2]), dd_mul(a[2], b[1]));
    c[1] = dd_sub(dd_mul(a[2], b[0]), dd_mul(a[0], b[2]));
    c[2] = dd_sub(dd_mul(a[0], b[1]), dd_mul(a[1], b[0]));
}
__host__ __device__ inline double dd_dot(const D2 a[3], const D2 b[3]) {
    return dd_add(dd_add(dd_mul(a[0], b[0]), dd_mul(a[1], b[1])), dd_mul(a[2], b[2])).hi;
}

// the pair geometry in units of h: relative vertices d, plane offset z, per edge (w, lo, hi); shared by the moment
// terms and the solid-angle sum so that both see the same sign of z and w
struct PairGeo {
    double d[9];
    double r2[3];              // squared vertex distances
    double z;
    double w[3], lo[3], hi[3];
    double det;                // d0 . (d1 x d2), exact to rounding where `refined`
    bool refined;
};

// exact (double-double) vertex-minus-query and edge vectors, formed where they are used to keep few registers live
__host__ __device__ inline void dd_vq(const double* __restrict__ fr, int k, const double q[3], D2 out[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = two_sum(fr[FR_V + 3 * k + c], -q[c]);
}
__host__ __device__ inline void dd_edge(const double* __restrict__ fr, int e, D2 out[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] = two_sum(fr[FR_V + 3 * ((e + 1) % 3) + c], -fr[FR_V + 3 * e + c]);
}

// z, det and per edge (w, lo, hi) of a pair, recomputed from exact vertex differences
struct RefinedGeo {
    double z, det;
    double w[3], lo[3], hi[3];
};
__host__ __device__ __attribute__((noinline)) RefinedGeo refine_geo(const double* __restrict__ fr, double qx, double qy,
                                                                    double qz, double ih, int kz, bool near_start0,
                                                                    bool near_start1, bool near_start2);

__host__ __device__ inline void pair_geo(const double* __restrict__ fr, double qx, double qy, double qz, double ih, PairGeo& g) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        g.d[3 * k] = (fr[FR_V + 3 * k] - qx) * ih;
        g.d[3 * k + 1] = (fr[FR_V + 3 * k + 1] - qy) * ih;
        g.d[3 * k + 2] = (fr[FR_V + 3 * k + 2] - qz) * ih;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) g.r2[k] = g.d[3 * k] * g.d[3 * k] + g.d[3 * k + 1] * g.d[3 * k + 1] + g.d[3 * k + 2] * g.d[3 * k + 2];
    // The plane offset from the vertex nearest the query and each edge's line distance from its nearer end: a
    // product taken from a far vertex carries an absolute rounding of u |d_far|, which near another vertex is a
    // large relative error of z and w; the faces meeting at that vertex would then see it at inconsistent
    // positions, and their solid angles (O(1) there, weighted O(1) in the moments) would no longer cancel.
    const int kz = (g.r2[0] <= g.r2[1]) ? (g.r2[0] <= g.r2[2] ? 0 : 2) : (g.r2[1] <= g.r2[2] ? 1 : 2);
    const double* dz = g.d + 3 * kz;
    g.z = fr[FR_N] * dz[0] + fr[FR_N + 1] * dz[1] + fr[FR_N + 2] * dz[2];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double* t = fr + FR_T + 3 * e;
        const double* mu = fr + FR_MU + 3 * e;
        const double* v = g.d + 3 * e;
        const double* q = g.d + 3 * ((e + 1) % 3);
        const double* wv = g.r2[e] <= g.r2[(e + 1) % 3] ? v : q;
        // both ends projected from their own rounded vertices: near a corner the two edges meeting there must see
        // the same vertex, or the foot-point angles of the two edges disagree at the rounding scale of w
        g.w[e] = mu[0] * wv[0] + mu[1] * wv[1] + mu[2] * wv[2];
        g.lo[e] = t[0] * v[0] + t[1] * v[1] + t[2] * v[2];
        g.hi[e] = t[0] * q[0] + t[1] * q[1] + t[2] * q[2];
    }
    // Where the foot point of the query is near an edge line (|w_e| below 1e-3 of the distance of the edge's nearer
    // vertex) the edge angles atan(a s / (|w| r)) are steep in w and in the chord ends lo, hi, and the rounded unit
    // vectors of the record leave w_e, lo, hi (and z) with an absolute error of u |d|: the faces meeting at the edge
    // see the query at inconsistent positions (z small too) or the face's own angles lose digits at a / |w| (z large).
    // On a sliver face the rounded normal is off by ~u x aspect ratio, so every w_e carries that error against a
    // small width of the face.  In these cases z, w_e, lo_e, hi_e are recomputed from exact vertex differences
    // (double-double): z = N.(v_k - q) / |N|, w_e = N.((v - q) x E_e) / (|N| |E_e|), lo_e = E_e.(v_e - q) / |E_e|,
    // hi_e = E_e.(v_(e+1) - q) / |E_e|, with N = (v1 - v0) x (v2 - v0) and E_e the edge vector.
    constexpr double TAU2 = 1e-6;
    bool refine = fr[FR_SLIVER] != 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) refine |= g.w[e] * g.w[e] < TAU2 * fmin(g.r2[e], g.r2[(e + 1) % 3]);
    g.refined = refine;
    if (refine) {
        const RefinedGeo r = refine_geo(fr, qx, qy, qz, ih, kz, g.r2[0] <= g.r2[1], g.r2[1] <= g.r2[2], g.r2[2] <= g.r2[0]);
        g.z = r.z;
        g.det = r.det;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            g.w[e] = r.w[e];
            g.lo[e] = r.lo[e];
            g.hi[e] = r.hi[e];
        }
    }
}

// z, w_e, lo_e, hi_e from exact vertex differences (see pair_geo); out of line: a rare path that would otherwise
// add its registers to every caller's.  near_start_e: the edge's start vertex is the nearer end.
__host__ __device__ __attribute__((noinline)) RefinedGeo refine_geo(const double* __restrict__ fr, double qx, double qy,
                                                                    double qz, double ih, int kz, bool near_start0,
                                                                    bool near_start1, bool near_start2) {
    RefinedGeo r;
    const double qv[3] = {qx, qy, qz};
    const bool near_start[3] = {near_start0, near_start1, near_start2};
    D2 N[3];
    {
        D2 E0[3], E2[3];
        dd_edge(fr, 0, E0);
        dd_edge(fr, 2, E2);
        dd_cross(E2, E0, N);                          // (v0 - v2) x (v1 - v0) = (v1 - v0) x (v2 - v0)
    }
    const double inl = ih / fr[FR_NLEN];
    {
        D2 D[3];
        dd_vq(fr, kz, qv, D);
        const double det = dd_dot(N, D);              // N.(v_k - q), the same for every k
        r.z = det * inl;
        r.det = det * (ih * ih * ih);
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        D2 E[3], D[3], C[3];
        dd_edge(fr, e, E);
        const double il = ih / fr[FR_LEN + e];
        dd_vq(fr, e, qv, D);
        r.lo[e] = dd_dot(E, D) * il;
        if (near_start[e]) {
            dd_cross(D, E, C);
            dd_vq(fr, (e + 1) % 3, qv, D);
        } else {
            dd_vq(fr, (e + 1) % 3, qv, D);
            dd_cross(D, E, C);
        }
        r.w[e] = dd_dot(N, C) * (inl / fr[FR_LEN + e]);
        r.hi[e] = dd_dot(E, D) * il;
    }
    return r;
}

// angle of the triangle at the foot point of the query: 2 pi when the foot is strictly inside (every w_e > 0),
// 0 when it is strictly outside one edge line, else the sum of the edge angles (foot on an edge line)
template <int FORM>
__host__ __device__ inline double foot_angle(const PairGeo& g) {
    if (g.w[0] > 0.0 && g.w[1] > 0.0 && g.w[2] > 0.0) return TWO_PI;
    if (g.w[0] < 0.0 || g.w[1] < 0.0 || g.w[2] < 0.0) return 0.0;
    double th = 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double b = fabs(g.w[e]);
        if (b > 0.0) th += sgn(g.w[e]) * angle2<FORM>(b * (g.hi[e] - g.lo[e]), b * b + g.lo[e] * g.hi[e]);
    }
    return th;
}

// angle subtended at the foot point by the edge piece [x1, x2] (b = |w| > 0): atan(x2/b) - atan(x1/b)
template <int FORM>
__host__ __device__ inline double piece_theta(double b, double x1, double x2) {
    return angle2<FORM>(b * (x2 - x1), b * b + x1 * x2);
}

// atan(a x2/(b r2)) - atan(a x1/(b r1)) for the piece [x1, x2] (r_i = sqrt(x_i^2 + c2), c2 = b^2 + a^2), with
// x2 r1 - x1 r2 formed without cancellation
template <int FORM>
__host__ __device__ inline double piece_phi(double a, double b, double c2, double x1, double r1, double x2, double r2) {
    double D;
    if (x1 >= 0.0 || x2 <= 0.0) {
        const double den = x2 * r1 + x1 * r2;
        D = den != 0.0 ? c2 * (x2 - x1) * (x2 + x1) / den : 0.0;
    } else {
        D = x2 * r1 - x1 * r2;
    }
    return angle2<FORM>(a * b * D, b * b * r1 * r2 + a * a * x1 * x2);
}

// distances of the three vertices from the query (units of h): the endpoint radii of every unclipped edge end
template <int FORM>
__host__ __device__ inline void vertex_radii(const PairGeo& g, double rv[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) rv[k] = xsqrt<FORM>(g.r2[k]);
}

// solid angle of the whole face seen from the query with the reference's convention (0 for a face whose plane
// contains the query): z S_f[r^-3] over the whole face = sgn(z) (Theta - sum_e sgn(w_e) dA_e), dA_e the angle of
// a s/(|w| r) across the edge (mesh3d/reference.py:127-150 at R = None, n = -3).  FP64 angles.
__host__ __device__ inline double omega_k(const PairGeo& g, const double rv[3]) {
    if (g.z == 0.0) return 0.0;
    const double a = fabs(g.z);
    double s = foot_angle<FP64>(g);
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double b = fabs(g.w[e]);
        if (b > 0.0) {
            const double c2 = b * b + a * a;
            s -= sgn(g.w[e]) * piece_phi<FP64>(a, b, c2, g.lo[e], rv[e], g.hi[e], rv[(e + 1) % 3]);
        }
    }
    return sgn(g.z) * s;
}

// the same solid angle for the mixed form: 2 atan2(det, den) of the corner vectors (van Oosterom and Strackee), one
// angle instead of three; where det and den are both below 1e-2 of their scale (the query near an edge of the face,
// where the formula's angle carries u / max(|det|, |den|) of rounding and the faces of the edge see the query at
// inconsistent positions) the form above is taken instead.  The FP64 form keeps
// the form above everywhere: this angle enters the moments with weights of up to ~1e3 (w6), and the corner-vector
// formula's rounding (~10 u of its scale) would then exceed 1e-12.
template <int FORM>
__host__ __device__ inline double omega_face(const PairGeo& g, const double rv[3]) {
    if (g.refined) {
        // exact triple product: 2 atan2(det, den) keeps full relative accuracy where den has no cancellation (den
        // at least la lb lc), in particular for the small solid angle of a sliver seen from afar, which the three-angle
        // form obtains as 2 pi minus angles of ~pi
        if (g.z == 0.0) return 0.0;
        const double* A = g.d;
        const double* B = g.d + 3;
        const double* C = g.d + 6;
        const double la = rv[0], lb = rv[1], lc = rv[2];
        const double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
        const double bc = B[0] * C[0] + B[1] * C[1] + B[2] * C[2];
        const double ca = C[0] * A[0] + C[1] * A[1] + C[2] * A[2];
        const double lll = la * lb * lc;
        const double den = lll + ab * lc + bc * la + ca * lb;
        if (den >= lll) return 2.0 * angle2<FP64>(g.det, den);
        return omega_k(g, rv);
    }
    if (FORM == MIXED) {
        if (g.z == 0.0) return 0.0;
        const double* A = g.d;
        const double* B = g.d + 3;
        const double* C = g.d + 6;
        const double det = A[0] * (B[1] * C[2] - B[2] * C[1]) + A[1] * (B[2] * C[0] - B[0] * C[2]) + A[2] * (B[0] * C[1] - B[1] * C[0]);
        const double la = rv[0], lb = rv[1], lc = rv[2];
        const double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
        const double bc = B[0] * C[0] + B[1] * C[1] + B[2] * C[2];
        const double ca = C[0] * A[0] + C[1] * A[1] + C[2] * A[2];
        const double den = la * lb * lc + ab * lc + bc * la + ca * lb;
        const double scale = la * lb * lc;
        if (fabs(det) > 1e-2 * scale || fabs(den) > 1e-2 * scale) return 2.0 * angle2<FORM>(det, den);
    }
    return omega_k(g, rv);
}

template <int FORM>
__host__ __device__ inline double face_solid_angle(const double* __restrict__ fr, double qx, double qy, double qz, double ih) {
    PairGeo g;
    pair_geo(fr, qx, qy, qz, ih, g);
    double rv[3];
    vertex_radii<FP64>(g, rv);
    return omega_face<FORM>(g, rv);
}

// One (query, face) pair: the 23 face accumulators in units of h (before 1/pi).  fr = face record, q = query,
// ih = 1/h.
//
// With the in-chord K chain split into its chord-integral part J_k (K_k = J_k + a^k K_0 for even k,
// K_k = J_k + a^k (a K_-1) for odd k) and the identities sum_e sgn(w_e) theta_e = Theta and
// Theta - sum_e sgn(w_e) phi_e = sgn(z) Omega_f over whole edges, the face integral becomes
//
//     S_f[g] = sum_e sum_m g_m/(m+2) w_e dJ_(m+2)  +  (G(R) - G_even(a)) sum_e sgn(w_e) theta_out,e
//              - G_odd(a) (sgn(z) Omega_f + sum_e sgn(w_e) phi_out,e),
//     z S_f[r^-3] = Omega_f + sgn(z) (sum_e sgn(w_e) phi_out,e - (a/R) sum_e sgn(w_e) theta_out,e),
//
// where theta_out, phi_out are the angles of the edge pieces outside the ball (zero for a face inside the support):
// no angle enters as a difference of nearly equal angles, so FP32 angles keep their relative accuracy.
template <int KID, int FORM>
__host__ __device__ inline void face_terms(const double* __restrict__ fr, double qx, double qy, double qz, double ih,
                                           double acc[NACC]) {
    using T = K3Tab<KID>;
    constexpr int DG = T::DEG;
    PairGeo g;
    pair_geo(fr, qx, qy, qz, ih, g);
    const double z = g.z, a = fabs(z), sz = sgn(z);
    const double nx = fr[FR_N], ny = fr[FR_N + 1], nz = fr[FR_N + 2];
    if (!(a < T::R[0])) return;                     // R[0] is the largest block radius
    double rv[3];
    vertex_radii<FORM>(g, rv);
    const double Om = omega_face<FORM>(g, rv);
    double apow[DG + 3];
    apow[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= DG + 2; ++k) apow[k] = apow[k - 1] * a;
    double sP = 0.0, sF = 0.0, zs0 = 0.0, zsV = 0.0;
    double l1x = 0.0, l1y = 0.0, l1z = 0.0, l2x = 0.0, l2y = 0.0, l2z = 0.0;
#pragma unroll
    for (int b = 0; b < T::NB; ++b) {
        const double R = T::R[b];
        if (!(a < R)) continue;
        double tP = 0.0, tF = 0.0, tpP = 0.0, tpF = 0.0, thO = 0.0, phO = 0.0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const double w = g.w[e], lo = g.lo[e], hi = g.hi[e];
            const double c2 = w * w + a * a;
            const double bw = fabs(w), sw = sgn(w);
            bool chord = false, clo = false, chi = false;
            double L = 0.0, l = 0.0, u = 0.0;
            if (c2 < R * R) {
                L = xsqrt<FORM>(R * R - c2);
                clo = lo < -L;
                chi = hi > L;
                l = clo ? -L : lo;
                u = chi ? L : hi;
                chord = l < u;
            }
            const double rlo = rv[e], rhi = rv[(e + 1) % 3];      // radii at the edge's vertices
            // angles of the pieces outside the ball
            if (bw > 0.0) {
                if (!chord) {
                    thO += sw * piece_theta<FORM>(bw, lo, hi);
                    if (a > 0.0) phO += sw * piece_phi<FORM>(a, bw, c2, lo, rlo, hi, rhi);
                } else {
                    if (clo) {
                        thO += sw * piece_theta<FORM>(bw, lo, -L);
                        if (a > 0.0) phO += sw * piece_phi<FORM>(a, bw, c2, lo, rlo, -L, R);
                    }
                    if (chi) {
                        thO += sw * piece_theta<FORM>(bw, L, hi);
                        if (a > 0.0) phO += sw * piece_phi<FORM>(a, bw, c2, L, R, hi, rhi);
                    }
                }
            }
            if (!chord) continue;
            const double rl = clo ? R : rlo;
            const double ru = chi ? R : rhi;
            // I_-1 = asinh(s/c): one log of F(u)/F(l), F(s) = s + r (= c^2/(r - s) for s < 0)
            double dIm1 = 0.0;
            if (c2 > 0.0) {
                const double Nu = u >= 0.0 ? u + ru : c2, Du = u >= 0.0 ? 1.0 : ru - u;
                const double Nl = l >= 0.0 ? l + rl : c2, Dl = l >= 0.0 ? 1.0 : rl - l;
                dIm1 = log((Nu * Dl) / (Du * Nl));
            }
            // chord integrals dI_k and the w-weighted chord parts wJ_k of the K chain, one unrolled pass
            double dI[DG + 1];
            double wJ[DG + 1];
            dI[0] = u - l;
            double pu = u, pl = l;
            wJ[0] = 0.0;
            if (DG >= 1) {
                pu *= ru;
                pl *= rl;
                dI[1] = 0.5 * ((pu - pl) + c2 * dIm1);
                wJ[1] = w * dIm1;
            }
#pragma unroll
            for (int k = 2; k <= DG; ++k) {
                pu *= ru;
                pl *= rl;
                dI[k] = ((pu - pl) + (double)k * c2 * dI[k - 2]) * (1.0 / (double)(k + 1));
                wJ[k] = w * dI[k - 2] + apow[2] * wJ[k - 2];
            }
            double eP = 0.0, eF = 0.0, epP = 0.0, epF = 0.0, lF = 0.0, lF2 = 0.0;
#pragma unroll
            for (int m = 0; m + 2 <= DG; ++m) {
                if (T::QP[b][m] != 0.0) eP += T::QP[b][m] * wJ[m + 2];
                if (T::QF[b][m] != 0.0) eF += T::QF[b][m] * wJ[m + 2];
                if (T::QpP[b][m] != 0.0) epP += T::QpP[b][m] * wJ[m + 2];
                if (T::QpF[b][m] != 0.0) epF += T::QpF[b][m] * wJ[m + 2];
            }
#pragma unroll
            for (int m = 0; m <= DG; ++m) {
                if (T::F[b][m] != 0.0) lF += T::F[b][m] * dI[m];
                if (T::F2[b][m] != 0.0) lF2 += T::F2[b][m] * dI[m];
            }
            tP += eP;
            tF += eF;
            tpP += epP;
            tpF += epF;
            const double* mu = fr + FR_MU + 3 * e;
            l1x += mu[0] * lF;
            l1y += mu[1] * lF;
            l1z += mu[2] * lF;
            l2x += mu[0] * lF2;
            l2y += mu[1] * lF2;
            l2z += mu[2] * lF2;
        }
        // G_even(a), G_odd(a) of each radial function
        double eP2 = 0.0, eF2 = 0.0, epP2 = 0.0, epF2 = 0.0, oP = 0.0, oF = 0.0, opP = 0.0, opF = 0.0;
#pragma unroll
        for (int m = 0; m + 2 <= DG; ++m) {
            const double am = apow[m + 2];
            if (m % 2 == 0) {
                if (T::QP[b][m] != 0.0) eP2 += T::QP[b][m] * am;
                if (T::QF[b][m] != 0.0) eF2 += T::QF[b][m] * am;
                if (T::QpP[b][m] != 0.0) epP2 += T::QpP[b][m] * am;
                if (T::QpF[b][m] != 0.0) epF2 += T::QpF[b][m] * am;
            } else {
                if (T::QP[b][m] != 0.0) oP += T::QP[b][m] * am;
                if (T::QF[b][m] != 0.0) oF += T::QF[b][m] * am;
                if (T::QpP[b][m] != 0.0) opP += T::QpP[b][m] * am;
                if (T::QpF[b][m] != 0.0) opF += T::QpF[b][m] * am;
            }
        }
        const double oph = sz * Om + phO;
        const double SP = tP + (T::GRP[b] - eP2) * thO - oP * oph;
        const double SF = tF + (T::GRF[b] - eF2) * thO - oF * oph;
        const double SpP = tpP + (T::GRpP[b] - epP2) * thO - opP * oph;
        const double SpF = tpF + (T::GRpF[b] - epF2) * thO - opF * oph;
        const double zS3 = Om + sz * (phO - (a / R) * thO);
        sP += SP;
        sF += SF;
        zs0 += z * SpP - T::MR[b] * zS3;
        zsV += z * SpF - T::MRF[b] * zS3;
    }
    const double n[3] = {nx, ny, nz};
    const double l1[3] = {l1x, l1y, l1z};
    const double l2[3] = {l2x, l2y, l2z};
    acc[0] += zs0;
    acc[1] += nx * sF;
    acc[2] += ny * sF;
    acc[3] += nz * sF;
    acc[4] += zsV;
    int k = 5;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) acc[k++] += n[i] * (z * n[j] * sF + l2[j]);
    acc[11] -= nx * sP;
    acc[12] -= ny * sP;
    acc[13] -= nz * sP;
    k = 14;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[k++] -= n[j] * (z * n[d] * sP + l1[d]);
}

// The six Hessian entries of baseline(3)[22:28] (fluxbridge/contract.py:41-46; mesh3d/reference.py:296-305 at
// alpha = 0): int_V d_i d_j W = sum_f n_i int_f y_j L(r) dS with L = W'(r)/r (the sphere part vanishes: W' = 0 at
// every block radius).  On a face y = z n + rho, so int_f y_j L = z n_j S_f[L] + sum_e mu_e,j int_e (W - W(R)) ds
// over the in-ball chord (the in-plane divergence theorem; W - W(R) vanishes on the arc).  S_f[L] is the face
// closed form above with the K-chain weights HW[k] at index k = m + 2 (L has the power r^-1 where W has r^1, whose
// K_1 = I_-1 + a^2 K_-1 the chain already holds); the edge part weights the same chord integrals dI_k with HW[k]
// (-G(R) at k = 0).  acc[0..5] += the face's (i, j) terms for (0,0) (0,1) (0,2) (1,1) (1,2) (2,2), units of h.
constexpr int NHESS = 6;

template <int KID, int FORM>
__host__ __device__ inline void face_hessian(const double* __restrict__ fr, double qx, double qy, double qz, double ih,
                                             double acc[NHESS]) {
    using T = K3Tab<KID>;
    constexpr int DG = T::DEG;
    PairGeo g;
    pair_geo(fr, qx, qy, qz, ih, g);
    const double z = g.z, a = fabs(z), sz = sgn(z);
    if (!(a < T::R[0])) return;
    double rv[3];
    vertex_radii<FORM>(g, rv);
    const double Om = omega_face<FORM>(g, rv);
    double apow[DG + 3];
    apow[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= DG + 2; ++k) apow[k] = apow[k - 1] * a;
    double sL = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
#pragma unroll
    for (int b = 0; b < T::NB; ++b) {
        const double R = T::R[b];
        if (!(a < R)) continue;
        double tL = 0.0, thO = 0.0, phO = 0.0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const double w = g.w[e], lo = g.lo[e], hi = g.hi[e];
            const double c2 = w * w + a * a;
            const double bw = fabs(w), sw = sgn(w);
            bool chord = false, clo = false, chi = false;
            double L = 0.0, l = 0.0, u = 0.0;
            if (c2 < R * R) {
                L = xsqrt<FORM>(R * R - c2);
                clo = lo < -L;
                chi = hi > L;
                l = clo ? -L : lo;
                u = chi ? L : hi;
                chord = l < u;
            }
            const double rlo = rv[e], rhi = rv[(e + 1) % 3];
            if (bw > 0.0) {
                if (!chord) {
                    thO += sw * piece_theta<FORM>(bw, lo, hi);
                    if (a > 0.0) phO += sw * piece_phi<FORM>(a, bw, c2, lo, rlo, hi, rhi);
                } else {
                    if (clo) {
                        thO += sw * piece_theta<FORM>(bw, lo, -L);
                        if (a > 0.0) phO += sw * piece_phi<FORM>(a, bw, c2, lo, rlo, -L, R);
                    }
                    if (chi) {
                        thO += sw * piece_theta<FORM>(bw, L, hi);
                        if (a > 0.0) phO += sw * piece_phi<FORM>(a, bw, c2, L, R, hi, rhi);
                    }
                }
            }
            if (!chord) continue;
            const double rl = clo ? R : rlo;
            const double ru = chi ? R : rhi;
            double dIm1 = 0.0;
            if (c2 > 0.0) {
                const double Nu = u >= 0.0 ? u + ru : c2, Du = u >= 0.0 ? 1.0 : ru - u;
                const double Nl = l >= 0.0 ? l + rl : c2, Dl = l >= 0.0 ? 1.0 : rl - l;
                dIm1 = log((Nu * Dl) / (Du * Nl));
            }
            double dI[DG + 1];
            double wJ[DG + 1];
            dI[0] = u - l;
            double pu = u, pl = l;
            wJ[0] = 0.0;
            if (DG >= 1) {
                pu *= ru;
                pl *= rl;
                dI[1] = 0.5 * ((pu - pl) + c2 * dIm1);
                wJ[1] = w * dIm1;
            }
#pragma unroll
            for (int k = 2; k <= DG; ++k) {
                pu *= ru;
                pl *= rl;
                dI[k] = ((pu - pl) + (double)k * c2 * dI[k - 2]) * (1.0 / (double)(k + 1));
                wJ[k] = w * dI[k - 2] + apow[2] * wJ[k - 2];
            }
            double eL = 0.0, lW = 0.0;
#pragma unroll
            for (int k = 0; k <= DG; ++k) {
                if (T::HW[b][k] != 0.0) {
                    if (k >= 1) eL += T::HW[b][k] * wJ[k];
                    lW += T::HW[b][k] * dI[k];
                }
            }
            tL += eL;
            const double* mu = fr + FR_MU + 3 * e;
            hx += mu[0] * lW;
            hy += mu[1] * lW;
            hz += mu[2] * lW;
        }
        double eL2 = 0.0, oL = 0.0;
#pragma unroll
        for (int k = 1; k <= DG; ++k) {
            if (T::HW[b][k] == 0.0) continue;
            if (k % 2 == 0) eL2 += T::HW[b][k] * apow[k];
            else oL += T::HW[b][k] * apow[k];
        }
        sL += tL + (T::GRL[b] - eL2) * thO - oL * (sz * Om + phO);
    }
    const double n[3] = {fr[FR_N], fr[FR_N + 1], fr[FR_N + 2]};
    const double hl[3] = {hx, hy, hz};
    int k = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) acc[k++] += n[i] * (z * n[j] * sL + hl[j]);
}

// Hessian accumulators -> the six entries (x 1/pi of the tables), physical units (h^-2) when `physical`
__host__ __device__ inline void finalize_hessian(const double acc[NHESS], double h, bool physical, double out[NHESS]) {
    const double s = physical ? INV_PI / (h * h) : INV_PI;
#pragma unroll
    for (int k = 0; k < NHESS; ++k) out[k] = acc[k] * s;
}

// accumulators + winding number -> the 22 channels of baseline(3)[:22] (fluxbridge/contract.py:41-46), physical
// units when `physical` (h^(|alpha| - |derivative|), fluxbridge/evaluate.py:76-78)
template <int KID>
__host__ __device__ inline void finalize22(const double acc[NACC], double chi, double h, bool physical, double out[NOUT]) {
    using T = K3Tab<KID>;
    const double M0 = acc[0] + 4.0 * PI * chi * T::MRSUM;
    const double V = acc[4] + 4.0 * PI * chi * T::MRFSUM;
    out[0] = M0;
    out[1] = acc[1];
    out[2] = acc[2];
    out[3] = acc[3];
    out[4] = acc[5] - V;
    out[5] = acc[6];
    out[6] = acc[7];
    out[7] = acc[8] - V;
    out[8] = acc[9];
    out[9] = acc[10] - V;
    out[10] = acc[11];
    out[11] = acc[12];
    out[12] = acc[13];
#pragma unroll
    for (int k = 0; k < 9; ++k) out[13 + k] = acc[14 + k] + ((k % 4) == 0 ? M0 : 0.0);
#pragma unroll
    for (int k = 0; k < NOUT; ++k) out[k] *= INV_PI;
    if (physical) {
        const double hh = h * h, ih = 1.0 / h;
        out[1] *= h;
        out[2] *= h;
        out[3] *= h;
#pragma unroll
        for (int k = 4; k < 10; ++k) out[k] *= hh;
        out[10] *= ih;
        out[11] *= ih;
        out[12] *= ih;
    }
}

// ------------------------------------------------------------------------------------------------ static grid
struct Grid3 {
    double lo[3];
    double cell;
    int n[3];
    const int32_t* cstart;      // [ncells+1] dilated (support_max) candidate lists, ascending face ids
    const int32_t* citems;
    const float4* crec;         // [2*items] per cell item: face box rounded outward to FP32 (lo.xyz, id), (hi.xyz, 0)
    const double* cchi;         // [ncells] winding number of face-free cells
    const int32_t* cwalk;       // [ncells] -1 for a face-free cell, else the index d of the cell's axis walk
    const uint32_t* cwork;      // [ceil(ncells / 32)] bit c % 32 of word c / 32 set when cell c lists candidate faces
                                // or has an axis walk (the cell sort's test, one small table instead of two large ones)
    const int32_t* wdir;        // [ndirty] dir | steps << 3 (dir 0..5 = +x -x +y -y +z -z)
    const int32_t* wstart;      // [ndirty+1] faces a ray of the walk can cross (CSR), ascending face ids
    const int32_t* witems;
    const float4* wrec;         // [2*witems] per walk item, rounded outward to FP32: (u lo, u hi, v lo, v hi),
                                // (axis lo, axis hi, id, 0) with (u, v) the two axes across the walk
    const double* frec;         // [F*FREC]
    const double* fbox;         // [F*6] lo xyz, hi xyz
    int nf;
};

__host__ __device__ inline bool cell_of(const Grid3& g, double px, double py, double pz, int c[3]) {
    const double p[3] = {px, py, pz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double f = floor((p[k] - g.lo[k]) / g.cell);
        if (!(f >= 0.0 && f < (double)g.n[k])) return false;
        c[k] = (int)f;
    }
    return true;
}

__host__ __device__ inline int64_t cell_id(const Grid3& g, int i, int j, int k) {
    return ((int64_t)i * g.n[1] + j) * g.n[2] + k;
}

// The axis walk of a cell with faces: direction (axis ax, sign s), the winding number where the walk ends (a face-free
// cell, or 0 outside the grid; an integer) and the range [it0, it1) of walk items (the faces overlapping the walk's
// column, FP32 box records in wrec).
struct Walk {
    int ax, au, av, s, it0, it1;
    double acc;
};
__host__ __device__ inline Walk walk_of(const Grid3& g, const int c0[3], int d) {
    Walk w;
    const int walk = g.wdir[d];
    const int dir = walk & 7, steps = walk >> 3;
    w.ax = dir >> 1;
    w.au = (w.ax + 1) % 3;
    w.av = (w.ax + 2) % 3;
    w.s = (dir & 1) ? -1 : 1;
    int ce[3] = {c0[0], c0[1], c0[2]};
    ce[w.ax] += w.s * steps;
    w.acc = (ce[w.ax] >= 0 && ce[w.ax] < g.n[w.ax]) ? g.cchi[cell_id(g, ce[0], ce[1], ce[2])] : 0.0;
    w.it0 = g.wstart[d];
    w.it1 = g.wstart[d + 1];
    return w;
}

// one walk item (record q0, q1) against the ray p + t s e_ax, t > 0 (Pa, Pu, Pv: the coordinates of p along ax and
// the two other axes au, av): 0 no crossing, +1 / -1 a crossing where the ray leaves / enters the solid, 2 a
// predicate within its rounding bound (the ray through or near an edge or a vertex of the face, or p on or within
// rounding of its plane)
__host__ __device__ inline int walk_item(const double* __restrict__ frec, double px, double py, double pz, double Pa,
                                         double Pu, double Pv, int au, int av, int s, const float4& q0, const float4& q1) {
    if (Pu < (double)q0.x || Pu > (double)q0.y || Pv < (double)q0.z || Pv > (double)q0.w) return 0;
    if (s > 0 ? (double)q1.y < Pa : (double)q1.x > Pa) return 0;
    const int f = bits_int(q1.z);
    const double* fr = frec + FREC * (int64_t)f;
    const double Au = fr[FR_V + au], Av = fr[FR_V + av];
    const double Bu = fr[FR_V + 3 + au], Bv = fr[FR_V + 3 + av];
    const double Cu = fr[FR_V + 6 + au], Cv = fr[FR_V + 6 + av];
    // projected point-in-triangle: three orientations with a static rounding bound
    int pos = 0, neg = 0, unk = 0;
    const double eu[3] = {Bu - Au, Cu - Bu, Au - Cu}, ev[3] = {Bv - Av, Cv - Bv, Av - Cv};
    const double ou[3] = {Pu - Au, Pu - Bu, Pu - Cu}, ov[3] = {Pv - Av, Pv - Bv, Pv - Cv};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double t1 = eu[q] * ov[q], t2 = ev[q] * ou[q];
        const double o = t1 - t2;
        const double bound = 1e-14 * (fabs(t1) + fabs(t2)) + 1e-300;
        if (o > bound) ++pos;
        else if (o < -bound) ++neg;
        else ++unk;
    }
    if (pos > 0 && neg > 0) return 0;            // certainly outside the projected triangle
    if (unk > 0) return 2;
    // side of p against the face plane: N . (p - A), N = (B - A) x (C - A)
    const double Ax = fr[FR_V], Ay = fr[FR_V + 1], Az = fr[FR_V + 2];
    const double e1x = fr[FR_V + 3] - Ax, e1y = fr[FR_V + 4] - Ay, e1z = fr[FR_V + 5] - Az;
    const double e2x = fr[FR_V + 6] - Ax, e2y = fr[FR_V + 7] - Ay, e2z = fr[FR_V + 8] - Az;
    const double dx = px - Ax, dy = py - Ay, dz = pz - Az;
    const double Nx = e1y * e2z - e1z * e2y, Ny = e1z * e2x - e1x * e2z, Nz = e1x * e2y - e1y * e2x;
    const double side = Nx * dx + Ny * dy + Nz * dz;
    const double mag = (fabs(e1y * e2z) + fabs(e1z * e2y)) * fabs(dx) + (fabs(e1z * e2x) + fabs(e1x * e2z)) * fabs(dy) +
                       (fabs(e1x * e2y) + fabs(e1y * e2x)) * fabs(dz);
    if (!(fabs(side) > 1e-13 * mag + 1e-300)) return 2;
    // N_ax has the sign of the projected orientation (counter-clockwise in (au, av) <=> N_ax > 0); the ray
    // p + t s e_ax meets the plane at t = -side / (s N_ax): ahead of p <=> side and s N_ax differ in sign
    const double sgnN = pos > 0 ? 1.0 : -1.0;
    if ((side > 0.0) == (s * sgnN > 0.0)) return 0;
    // the ray leaves the solid where s N_ax > 0 (outward normal along the ray)
    return (s * sgnN > 0.0) ? 1 : -1;
}

// winding number at p from the axis walk of its cell: the winding number where the walk ends plus the signed
// crossings of the ray from p with the faces that overlap the walk's column.  Every such face lies inside the walked
// range (a face reaching the end cell would make it a cell with faces), so each crossing ahead of p is counted once.
// Returns false when a predicate is within its rounding bound (the caller then takes the exact per-face sum).  All
// terms are integers, so the sum is exact in any order.
__host__ __device__ inline bool ray_winding(const Grid3& g, double px, double py, double pz, const int c0[3], double* chi) {
    const int64_t cid0 = cell_id(g, c0[0], c0[1], c0[2]);
    const int d = g.cwalk[cid0];
    if (d < 0) {
        *chi = g.cchi[cid0];
        return true;
    }
    const Walk w = walk_of(g, c0, d);
    const double p[3] = {px, py, pz};
    const double Pa = p[w.ax], Pu = p[w.au], Pv = p[w.av];
    double acc = w.acc;
    for (int it0 = w.it0; it0 < w.it1; it0 += 4) {
        // four walk items per step: records loaded together, then the conservative FP32 box filter and the exact
        // tests on the face's vertices
        float4 r0[4], r1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = it0 + k < w.it1 ? it0 + k : w.it1 - 1;
            r0[k] = g.wrec[2 * (int64_t)j];
            r1[k] = g.wrec[2 * (int64_t)j + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (it0 + k >= w.it1) break;
            const int r = walk_item(g.frec, px, py, pz, Pa, Pu, Pv, w.au, w.av, w.s, r0[k], r1[k]);
            if (r == 2) return false;
            acc += (double)r;
        }
    }
    *chi = acc;
    return true;
}

}  // namespace m3
