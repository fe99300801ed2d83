// Double-double arithmetic (value = hi + lo, ~106 significant bits) from the published error-free transformations:
// Knuth's TwoSum and the FMA-based TwoProd.  Used where a closed form cancels more than FP64 can carry.
#pragma once
#include <hip/hip_runtime.h>

struct dd {
    double hi, lo;
};

__host__ __device__ __forceinline__ dd dd_of(double a) { return dd{a, 0.0}; }

__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return dd{s, e};
}

__host__ __device__ __forceinline__ dd quick_two_sum(double a, double b) {
#pragma clang fp contract(off)
    const double s = a + b;
    return dd{s, b - (s - a)};
}

__host__ __device__ __forceinline__ dd two_prod(double a, double b) {
    const double p = a * b;
    return dd{p, fma(a, b, -p)};
}

__host__ __device__ __forceinline__ dd dd_add(dd x, dd y) {
#pragma clang fp contract(off)
    dd s = two_sum(x.hi, y.hi);
    dd t = two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

__host__ __device__ __forceinline__ dd dd_neg(dd x) { return dd{-x.hi, -x.lo}; }
__host__ __device__ __forceinline__ dd dd_sub(dd x, dd y) { return dd_add(x, dd_neg(y)); }

__host__ __device__ __forceinline__ dd dd_mul(dd x, dd y) {
    dd p = two_prod(x.hi, y.hi);
    p.lo += x.hi * y.lo + x.lo * y.hi;
    return quick_two_sum(p.hi, p.lo);
}

__host__ __device__ __forceinline__ dd dd_mul_d(dd x, double y) {
    dd p = two_prod(x.hi, y);
    p.lo += x.lo * y;
    return quick_two_sum(p.hi, p.lo);
}

// x / y: one FP64 quotient estimate and one correction from the exact residual
__host__ __device__ __forceinline__ dd dd_div(dd x, dd y) {
    const double q1 = x.hi / y.hi;
    dd r = dd_sub(x, dd_mul_d(y, q1));
    const double q2 = r.hi / y.hi;
    r = dd_sub(r, dd_mul_d(y, q2));
    const double q3 = r.hi / y.hi;
    dd q = quick_two_sum(q1, q2);
    return dd_add(q, dd_of(q3));
}
