// C ABI plumbing: error state, version, plan upload.
#include <cstring>

#include "wsb_internal.h"

namespace {
thread_local std::string g_error;
}

namespace wsb {
int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}
}  // namespace wsb

extern "C" int wsb_last_error(char* buf, size_t n) {
    if (buf && n) {
        const size_t k = std::min(n - 1, g_error.size());
        std::memcpy(buf, g_error.data(), k);
        buf[k] = 0;
    }
    return (int)g_error.size();
}

extern "C" int wsb_version(void) { return 1; }

namespace {
template <class T>
int up(T** dst, const T* src, size_t count) {
    WSB_HIP(hipMalloc(dst, std::max<size_t>(count, 1) * sizeof(T)));
    if (count) WSB_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return WSB_OK;
}
}  // namespace

extern "C" int wsb_plan_create(int kernel_id, const double* radii, int nr, const int32_t* cn, const double* cc, int ncoef,
                               const int32_t* E, int nE, const int32_t* V, const double* mR, int nV, wsb_plan** out) {
    WSB_REQUIRE(out, WSB_EINVAL, "null output");
    WSB_REQUIRE(kernel_id >= 0 && kernel_id <= 3, WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
    WSB_REQUIRE(nr >= 1 && ncoef >= 0 && nE >= 0 && nV >= 0, WSB_EINVAL, "bad plan sizes");
    for (int t = 0; t < nE; ++t) {
        const int32_t* e = E + 7 * t;
        WSB_REQUIRE(e[0] >= 0 && e[0] < 12 && (e[1] == 0 || e[1] == 1) && e[2] >= 0 && e[3] >= 0 && e[4] >= 0 && e[4] < nr &&
                        e[5] >= 0 && e[6] <= ncoef && e[5] <= e[6],
                    WSB_EINVAL, "edge term %d out of range", t);
    }
    for (int t = 0; t < nV; ++t) {
        const int32_t* v = V + 4 * t;
        WSB_REQUIRE(v[0] >= 0 && v[0] < 12 && v[1] >= 0 && v[1] < nr && v[2] >= 0 && v[3] <= ncoef && v[2] <= v[3],
                    WSB_EINVAL, "value term %d out of range", t);
    }
    for (int k = 0; k < ncoef; ++k) WSB_REQUIRE(cn[k] >= 0 && cn[k] < 64, WSB_EINVAL, "coefficient power out of range");
    auto* p = new wsb_plan();
    WSB_HIP(hipGetDevice(&p->device));
    p->kernel_id = kernel_id;
    p->nr = nr;
    p->ncoef = ncoef;
    p->nE = nE;
    p->nV = nV;
    int st = up(&p->radii, radii, nr);
    if (!st) st = up(&p->cn, cn, ncoef);
    if (!st) st = up(&p->cc, cc, ncoef);
    if (!st) st = up(&p->E, E, 7 * (size_t)nE);
    if (!st) st = up(&p->V, V, 4 * (size_t)nV);
    if (!st) st = up(&p->mR, mR, nV);
    if (st) {
        wsb_plan_destroy(p);
        return st;
    }
    *out = p;
    return WSB_OK;
}

extern "C" int wsb_plan_destroy(wsb_plan* p) {
    if (!p) return WSB_OK;
    for (void* q : {(void*)p->radii, (void*)p->cn, (void*)p->cc, (void*)p->E, (void*)p->V, (void*)p->mR})
        if (q) (void)hipFree(q);
    delete p;
    return WSB_OK;
}
