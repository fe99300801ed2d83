// Solid-sphere closed form lambda_sph(R, d) in FP64, one thread per (R, d) pair (units of h).
//   lambda = sum_j [ 4 M3(a1) + ( -dM4 + 2 D dM3 - d (2R + d) dM2 ) / D ],  D = R + d,
//   M_k(r) = sum_n c_n r^(n+k)/(n+k), a1 = min(R_j, max(0, -d)), dM_k = M_k(min(R_j, 2R + d)) - M_k(|d|).
// The exact derivation (spherical caps of the support sphere) and its tables are in sphere.py / tables.inc.
#include "tables.inc"
#include "wsb_internal.h"

namespace {

template <int KID, int K>
__device__ __forceinline__ double Mk(int b, double r) {
    using T = STab<KID>;
    double acc = 0.0;
#pragma unroll
    for (int n = T::DEG; n >= 0; --n) {
        const double c = K == 2 ? T::M2[b][n] : (K == 3 ? T::M3[b][n] : T::M4[b][n]);
        acc = fma(acc, r, c);
    }
    double rk = r * r;
    if (K >= 3) rk *= r;
    if (K >= 4) rk *= r;
    return acc * rk;
}

template <int KID>
__global__ __launch_bounds__(256) void sphere_kernel(const double* Rv, const double* dv, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using T = STab<KID>;
    const double R = Rv[i], d = dv[i];
    const double D = R + d;
    double tot = 0.0;
#pragma unroll
    for (int b = 0; b < T::NB; ++b) {
        const double Rj = T::R[b];
        const double a1 = fmin(Rj, fmax(0.0, -d));
        if (a1 > 0.0) tot += 4.0 * Mk<KID, 3>(b, a1);
        const double a = fabs(d), hi = fmin(Rj, 2.0 * R + d);
        if (a < hi && D > 0.0) {
            const double dM2 = Mk<KID, 2>(b, hi) - Mk<KID, 2>(b, a);
            const double dM3 = Mk<KID, 3>(b, hi) - Mk<KID, 3>(b, a);
            const double dM4 = Mk<KID, 4>(b, hi) - Mk<KID, 4>(b, a);
            tot += 2.0 * dM3 - (dM4 + d * (2.0 * R + d) * dM2) / D;
        }
    }
    out[i] = (R > 0.0 && d >= -R) ? tot : __builtin_nan("");
}

}  // namespace

extern "C" int wsb_sphere_closed_form(int kernel_id, const double* R, const double* d, int n, double* out, void* stream) {
    WSB_REQUIRE(n >= 0, WSB_EINVAL, "negative count");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(R && d && out, WSB_EINVAL, "null buffer");
    hipStream_t s = wsb::as_stream(stream);
    const dim3 g((n + 255) / 256), b(256);
    switch (kernel_id) {
        case WSB_CUBIC: hipLaunchKernelGGL(sphere_kernel<WSB_CUBIC>, g, b, 0, s, R, d, n, out); break;
        case WSB_W2: hipLaunchKernelGGL(sphere_kernel<WSB_W2>, g, b, 0, s, R, d, n, out); break;
        case WSB_W4: hipLaunchKernelGGL(sphere_kernel<WSB_W4>, g, b, 0, s, R, d, n, out); break;
        case WSB_W6: hipLaunchKernelGGL(sphere_kernel<WSB_W6>, g, b, 0, s, R, d, n, out); break;
        default: return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}
