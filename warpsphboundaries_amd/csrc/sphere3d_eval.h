// Per-pair evaluation of the solid-sphere closed form (host and device), see sphere3d.hip.
#pragma once
#include "dd.h"
#include "tables.inc"

template <int KID, int K>
__host__ __device__ __forceinline__ dd sphere_Mk(int b, dd r) {
    using T = STab<KID>;
    dd acc = dd_of(0.0);
#pragma unroll
    for (int n = T::DEG; n >= 0; --n) {
        const double ch = K == 2 ? T::M2H[b][n] : (K == 3 ? T::M3H[b][n] : T::M4H[b][n]);
        const double cl = K == 2 ? T::M2L[b][n] : (K == 3 ? T::M3L[b][n] : T::M4L[b][n]);
        acc = dd_add(dd_mul(acc, r), dd{ch, cl});
    }
    dd rk = dd_mul(r, r);
    if (K >= 3) rk = dd_mul(rk, r);
    if (K >= 4) rk = dd_mul(rk, r);
    return dd_mul(acc, rk);
}

template <int KID>
__host__ __device__ __forceinline__ double sphere_lambda(double R, double d) {
    using T = STab<KID>;
    const dd D = two_sum(R, d);
    const dd b0 = dd_add(two_sum(R, R), dd_of(d));           // 2R + d
    dd tot = dd_of(0.0);
#pragma unroll
    for (int b = 0; b < T::NB; ++b) {
        const double Rj = T::R[b];
        const double a1 = fmin(Rj, fmax(0.0, -d));
        if (a1 > 0.0) tot = dd_add(tot, dd_mul_d(sphere_Mk<KID, 3>(b, dd_of(a1)), 4.0));
        const dd a = dd_of(fabs(d));
        const dd hi = (b0.hi > Rj || (b0.hi == Rj && b0.lo > 0.0)) ? dd_of(Rj) : b0;
        const bool open = (a.hi < hi.hi || (a.hi == hi.hi && a.lo < hi.lo)) && (D.hi > 0.0);
        if (open) {
            const dd dM2 = dd_sub(sphere_Mk<KID, 2>(b, hi), sphere_Mk<KID, 2>(b, a));
            const dd dM3 = dd_sub(sphere_Mk<KID, 3>(b, hi), sphere_Mk<KID, 3>(b, a));
            const dd dM4 = dd_sub(sphere_Mk<KID, 4>(b, hi), sphere_Mk<KID, 4>(b, a));
            const dd num = dd_add(dM4, dd_mul(dd_mul_d(b0, d), dM2));
            tot = dd_add(tot, dd_sub(dd_mul_d(dM3, 2.0), dd_div(num, D)));
        }
    }
    return (R > 0.0 && d >= -R) ? tot.hi + tot.lo : __builtin_nan("");
}
