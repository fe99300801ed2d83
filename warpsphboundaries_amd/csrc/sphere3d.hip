// Solid-sphere closed form lambda_sph(R, d), one thread per (R, d) pair (units of h):
//   lambda = sum_j [ 4 M3(a1) + 2 dM3 - ( dM4 + d (2R + d) dM2 ) / D ],  D = R + d,
//   M_k(r) = sum_n c_n r^(n+k)/(n+k), a1 = min(R_j, max(0, -d)), dM_k = M_k(min(R_j, 2R + d)) - M_k(|d|).
// The bracket cancels by about 1/D (query near the centre of a ball) times the size of the monomial coefficients
// (up to 1e4 for w6), so it is carried in double-double arithmetic with the exact coefficients split into hi + lo:
// the result is FP64-accurate.  The derivation (spherical caps of the support sphere) is in sphere.py.
#include "sphere3d_eval.h"
#include "wsb_internal.h"

namespace {

template <int KID>
__global__ __launch_bounds__(256) void sphere_kernel(const double* Rv, const double* dv, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = sphere_lambda<KID>(Rv[i], dv[i]);
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
