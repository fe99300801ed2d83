// Twelve-moment evaluation of 2-D surface representations: the plain FP64 path (pair list + plan
// interpreter + row sums) and the particle-tile path (fused pair evaluation and row sums).
#include <hipcub/hipcub.hpp>

#include "edge2d_device.h"

namespace {

constexpr int TILE = 256;          // rows per workgroup of the tile kernel (4 wavefronts)

inline size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// ------------------------------------------------------------------------------------------------ plain path
// scratch: off[n] | pair_q[cap] | pair_e[cap] | c[cap*12] | scan temp     (cap = n*K)
struct PlainLayout {
    size_t off, pq, pe, c, tmp, tmp_bytes, total;
};

PlainLayout plain_layout(int n, int K) {
    PlainLayout L{};
    size_t tmp = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const int32_t*)nullptr, (int32_t*)nullptr, n > 0 ? n : 1);
    const size_t cap = (size_t)n * (size_t)K;
    L.off = 0;
    L.pq = align_up(L.off + sizeof(int32_t) * (size_t)n);
    L.pe = align_up(L.pq + sizeof(int32_t) * cap);
    L.c = align_up(L.pe + sizeof(int32_t) * cap);
    L.tmp = align_up(L.c + sizeof(double) * 12 * cap);
    L.tmp_bytes = tmp;
    L.total = align_up(L.tmp + tmp);
    return L;
}

// every row writes its pairs (row, edge) at its offset: the CSR pair list of the fixed slots
__global__ __launch_bounds__(256) void rows_to_pairs_kernel(const int32_t* slots, const int32_t* counts, const int32_t* off,
                                                            int n, int K, int32_t* pq, int32_t* pe) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int c = counts[r], o = off[r];
    for (int s = 0; s < c; ++s) {
        pq[o + s] = r;
        pe[o + s] = slots[(int64_t)r * K + s];
    }
}

struct PlainPairArgs {
    plain::Plan plan;
    const int32_t* pq;
    const int32_t* pe;
    int P;                      // >= 0: explicit pair count; < 0: total from off/counts of the last row
    const int32_t* off;
    const int32_t* counts;
    int n;                      // rows of off/counts (P < 0)
    int nq, ne;                 // bounds of the query and edge indices
    const double* pos;
    const double* sup;
    const double* V;
    const int32_t* E;
    double* c;                  // [P*12]
};

// one thread per (query, edge) pair; grid-stride so that a device-side pair count needs no host read
__global__ __launch_bounds__(256) void plain_pairs_kernel(PlainPairArgs a) {
    const int P = a.P >= 0 ? a.P : (a.n > 0 ? a.off[a.n - 1] + a.counts[a.n - 1] : 0);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < P; j += gridDim.x * blockDim.x) {
        double ch[wsb::NCH];
        for (int k = 0; k < wsb::NCH; ++k) ch[k] = 0.0;
        const int e = a.pe[j], q = a.pq[j];
        if (e >= 0 && e < a.ne && q >= 0 && q < a.nq) {
            const double h = a.sup[q];
            const double x = a.pos[2 * q], y = a.pos[2 * q + 1];
            const int v0 = a.E[2 * e], v1 = a.E[2 * e + 1];
            const double px = (a.V[2 * v0] - x) / h, py = (a.V[2 * v0 + 1] - y) / h;
            const double qx = (a.V[2 * v1] - x) / h, qy = (a.V[2 * v1 + 1] - y) / h;
            plain::pair_channels(a.plan, px, py, qx, qy, ch);
        }
        for (int k = 0; k < wsb::NCH; ++k) a.c[(int64_t)j * wsb::NCH + k] = ch[k];
    }
}

struct RowsArgs {
    const double* c;
    const int32_t* off;
    const int32_t* counts;
    const double* ind;
    const double* lsup;
    int n;
    double plane[12];
    const double* rot4;
    int flags;
    double* out;
};

// per row: sum of its pairs in slot order, then indicator x plane vector (reference scene/fused.py:56-154)
__global__ __launch_bounds__(256) void plain_rows_kernel(RowsArgs a) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n) return;
    double acc[wsb::NCH];
    for (int k = 0; k < wsb::NCH; ++k) acc[k] = 0.0;
    const int o = a.off[r], c = a.counts[r];
    for (int s = 0; s < c; ++s)
        for (int k = 0; k < wsb::NCH; ++k) acc[k] += a.c[(int64_t)(o + s) * wsb::NCH + k];
    finalize_row(acc, a.ind[r], a.lsup[r], a.plane, a.rot4, a.flags, a.out + (int64_t)r * wsb::NCH);
}

// ------------------------------------------------------------------------------------------------ tile path
struct TileArgs {
    const int32_t* order;       // nullable: tile position i serves row order[i] (rows sorted by cell)
    const int32_t* slots;
    const int32_t* counts;
    int K;
    const double* ind;
    const double* lpos;
    const double* lsup;
    int n;
    const double4* erec;
    const double4* etan;
    int ne;
    const double* rot4;
    int flags;
    double* out;
};

// A workgroup owns TILE consecutive rows.  Their pairs are packed densely over the lanes (pair j of the tile
// -> its row by a binary search of the row offsets in LDS), every lane evaluates one pair with the closed
// form, the twelve channels go through LDS and each row's lane adds its own pairs in slot order: the row sums
// are deterministic and independent of how pairs fall into chunks.
template <int KID, int FORM>
__global__ __launch_bounds__(TILE) void tile12_kernel(TileArgs a) {
    using BlockScan = hipcub::BlockScan<int, TILE>;
    __shared__ typename BlockScan::TempStorage scan_tmp;
    __shared__ int s_off[TILE + 1];
    __shared__ double s_px[TILE], s_py[TILE], s_h[TILE], s_ih[TILE];
    __shared__ double s_res[wsb::NCH][TILE];
    __shared__ int s_row[TILE];
    const int tid = threadIdx.x;
    const int pos = blockIdx.x * TILE + tid;
    const bool live = pos < a.n;
    const int row = live ? (a.order ? a.order[pos] : pos) : 0;
    s_row[tid] = row;
    const int cnt = live ? min(max(a.counts[row], 0), a.K) : 0;
    double h = 1.0;
    if (live) {
        s_px[tid] = a.lpos[2 * row];
        s_py[tid] = a.lpos[2 * row + 1];
        h = a.lsup[row];
        s_h[tid] = h;
        s_ih[tid] = 1.0 / h;
    }
    int off, total;
    BlockScan(scan_tmp).ExclusiveSum(cnt, off, total);
    s_off[tid] = off;
    if (tid == 0) s_off[TILE] = total;
    __syncthreads();
    double acc[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) acc[k] = 0.0;
    for (int base = 0; base < total; base += TILE) {
        const int j = base + tid;
        double c[wsb::NCH];
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) c[k] = 0.0;
        if (j < total) {
            int lo = 0, hi = TILE;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= j) lo = mid;
                else hi = mid;
            }
            const int r = lo;
            const int e = a.slots[(int64_t)s_row[r] * a.K + (j - s_off[r])];
            if (e >= 0 && e < a.ne) {
                const double x = s_px[r], y = s_py[r];
                const double4 ab = a.erec[e];
                const double4 tn = a.etan[e];
                if constexpr (KID >= WSB_W2P5)
                    chord::pair_grad<KID>(ab.x - x, ab.y - y, ab.z - x, ab.w - y, s_h[r], s_ih[r], tn.x, tn.y, tn.z, c);
                else
                    closed::pair_channels<KID, FORM>(ab.x - x, ab.y - y, ab.z - x, ab.w - y, s_h[r], s_ih[r], tn.x, tn.y,
                                                     tn.z, c);
            }
        }
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) s_res[k][tid] = c[k];
        __syncthreads();
        const int b0 = max(off, base), b1 = min(off + cnt, base + TILE);
        for (int jj = b0; jj < b1; ++jj) {
#pragma unroll
            for (int k = 0; k < wsb::NCH; ++k) acc[k] += s_res[k][jj - base];
        }
        __syncthreads();
    }
    // the W^5 kernels carry their gradient channels only (the whole-support gradient is zero: no indicator term)
    const int flags = KID >= WSB_W2P5 ? (a.flags | WSB_OUT_NO_INDICATOR) : a.flags;
    if (live) finalize_row(acc, a.ind[row], h, plane_of<KID>(), a.rot4, flags, a.out + (int64_t)row * wsb::NCH);
}

template <int KID, int FORM>
int launch_tile(const TileArgs& a, hipStream_t s) {
    hipLaunchKernelGGL((tile12_kernel<KID, FORM>), dim3((a.n + TILE - 1) / TILE), dim3(TILE), 0, s, a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

template <int FORM>
int dispatch_tile(int kid, const TileArgs& a, hipStream_t s) {
    switch (kid) {
        case WSB_CUBIC: return launch_tile<WSB_CUBIC, FORM>(a, s);
        case WSB_W2: return launch_tile<WSB_W2, FORM>(a, s);
        case WSB_W4: return launch_tile<WSB_W4, FORM>(a, s);
        case WSB_W6: return launch_tile<WSB_W6, FORM>(a, s);
        case WSB_CONE: return launch_tile<WSB_CONE, FORM>(a, s);
        case WSB_LW2: return launch_tile<WSB_LW2, FORM>(a, s);
        case WSB_LW4: return launch_tile<WSB_LW4, FORM>(a, s);
        case WSB_W2P5: return launch_tile<WSB_W2P5, closed::FP64>(a, s);
        case WSB_W4P5: return launch_tile<WSB_W4P5, closed::FP64>(a, s);
    }
    return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kid);
}

// ------------------------------------------------------------------------------------------------ pair list path
struct PairArgs {
    const int32_t* pq;
    const int32_t* pe;
    int P, nq, ne;
    const double* pos;
    const double* sup;
    const double* V;
    const int32_t* E;
    double* out;
};

template <int KID, int FORM>
__global__ __launch_bounds__(256) void pairs12_kernel(PairArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.P) return;
    double c[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) c[k] = 0.0;
    const int e = a.pe[j], q = a.pq[j];
    if (e >= 0 && e < a.ne && q >= 0 && q < a.nq) {
        const double h = a.sup[q], x = a.pos[2 * q], y = a.pos[2 * q + 1];
        const int v0 = a.E[2 * e], v1 = a.E[2 * e + 1];
        const double ax = a.V[2 * v0], ay = a.V[2 * v0 + 1], bx = a.V[2 * v1], by = a.V[2 * v1 + 1];
        const double dx = bx - ax, dy = by - ay;
        const double len = sqrt(dx * dx + dy * dy);
        if (len > 0.0) closed::pair_channels<KID, FORM>(ax - x, ay - y, bx - x, by - y, h, 1.0 / h, dx / len, dy / len, 1.0 / len, c);
    }
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) a.out[(int64_t)j * wsb::NCH + k] = c[k];
}

template <int FORM>
int dispatch_pairs(int kid, const PairArgs& a, hipStream_t s) {
    const dim3 g((a.P + 255) / 256), b(256);
    switch (kid) {
        case WSB_CUBIC: hipLaunchKernelGGL((pairs12_kernel<WSB_CUBIC, FORM>), g, b, 0, s, a); break;
        case WSB_W2: hipLaunchKernelGGL((pairs12_kernel<WSB_W2, FORM>), g, b, 0, s, a); break;
        case WSB_W4: hipLaunchKernelGGL((pairs12_kernel<WSB_W4, FORM>), g, b, 0, s, a); break;
        case WSB_W6: hipLaunchKernelGGL((pairs12_kernel<WSB_W6, FORM>), g, b, 0, s, a); break;
        case WSB_CONE: hipLaunchKernelGGL((pairs12_kernel<WSB_CONE, FORM>), g, b, 0, s, a); break;
        case WSB_LW2: hipLaunchKernelGGL((pairs12_kernel<WSB_LW2, FORM>), g, b, 0, s, a); break;
        case WSB_LW4: hipLaunchKernelGGL((pairs12_kernel<WSB_LW4, FORM>), g, b, 0, s, a); break;
        default: return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kid);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

void plane_host(int kid, double* out) {
    static const double* const planes[] = {TTab<0>::PLANE, TTab<1>::PLANE, TTab<2>::PLANE, TTab<3>::PLANE, TTab<4>::PLANE,
                                           TTab<5>::PLANE, TTab<6>::PLANE, TTab<7>::PLANE, TTab<8>::PLANE};
    const double* p = planes[kid];
    for (int k = 0; k < 12; ++k) out[k] = p[k];
}

plain::Plan plan_view(const wsb_plan* p) {
    return plain::Plan{p->radii, p->cn, p->cc, p->E, p->V, p->mR, p->nE, p->nV};
}

int grid_for(int device) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    return cus * 8;
}

}  // namespace

extern "C" size_t wsb_moments2d_scratch_bytes(int n, int K) {
    if (n <= 0 || K <= 0) return 0;
    return plain_layout(n, K).total;
}

extern "C" int wsb_moments2d(const wsb_geom2d* g, const wsb_plan* plan, int kernel_id, int form, const int32_t* order,
                             const int32_t* slots, const int32_t* counts, int K, const double* ind, const double* lpos,
                             const double* lsup, int n, const double* rot4, int flags, double* out, void* scratch,
                             size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(g, WSB_EINVAL, "null geometry");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes n=%d K=%d", n, K);
    WSB_REQUIRE(kernel_id >= 0 && kernel_id <= WSB_W4P5, WSB_ENOTSUP, "unknown kernel id %d", kernel_id);
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD) || rot4, WSB_EINVAL, "WSB_OUT_WORLD needs rot4");
    WSB_REQUIRE(kernel_id < WSB_W2P5 || form != WSB_FORM_PLAIN, WSB_ENOTSUP,
                "the W^5 kernels have no plan interpreter (their monomial plan cancels to 1e-3)");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && counts && ind && lpos && lsup && out, WSB_EINVAL, "null buffer");
    hipStream_t s = wsb::as_stream(stream);
    if (form == WSB_FORM_PLAIN) {
        WSB_REQUIRE(plan, WSB_EINVAL, "the plain form needs a plan");
        WSB_REQUIRE(plan->kernel_id == kernel_id, WSB_EINVAL, "plan kernel %d != kernel %d", plan->kernel_id, kernel_id);
        const PlainLayout L = plain_layout(n, K);
        WSB_REQUIRE(scratch && scratch_bytes >= L.total, WSB_ENOMEM, "plain form needs %zu scratch bytes (got %zu)", L.total,
                    scratch_bytes);
        char* base = static_cast<char*>(scratch);
        int32_t* off = reinterpret_cast<int32_t*>(base + L.off);
        int32_t* pq = reinterpret_cast<int32_t*>(base + L.pq);
        int32_t* pe = reinterpret_cast<int32_t*>(base + L.pe);
        double* c = reinterpret_cast<double*>(base + L.c);
        size_t tmp = L.tmp_bytes;
        WSB_HIP(hipcub::DeviceScan::ExclusiveSum(base + L.tmp, tmp, counts, off, n, s));
        hipLaunchKernelGGL(rows_to_pairs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, slots, counts, off, n, K, pq, pe);
        WSB_HIP(hipGetLastError());
        PlainPairArgs pa{plan_view(plan), pq, pe, -1, off, counts, n, n, g->ne, lpos, lsup, g->verts, g->edges, c};
        hipLaunchKernelGGL(plain_pairs_kernel, dim3(grid_for(g->device)), dim3(256), 0, s, pa);
        WSB_HIP(hipGetLastError());
        RowsArgs ra{c, off, counts, ind, lsup, n, {0}, rot4, flags, out};
        plane_host(kernel_id, ra.plane);
        hipLaunchKernelGGL(plain_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, s, ra);
        WSB_HIP(hipGetLastError());
        return WSB_OK;
    }
    TileArgs ta{order, slots, counts, K, ind, lpos, lsup, n, g->erec, g->etan, g->ne, rot4, flags, out};
    if (form == WSB_FORM_FP64) return dispatch_tile<closed::FP64>(kernel_id, ta, s);
    if (form == WSB_FORM_MIXED) return dispatch_tile<closed::MIXED>(kernel_id, ta, s);
    return wsb::set_error(WSB_ENOTSUP, "unknown form %d", form);
}

extern "C" int wsb_edge_channels2d(const wsb_plan* plan, int kernel_id, int form, const int32_t* pair_q, const int32_t* pair_e,
                                   int P, const double* pos, const double* sup, int nq, const double* verts,
                                   const int32_t* edges, int ne, double* out, void* stream) {
    WSB_REQUIRE(P >= 0, WSB_EINVAL, "negative pair count");
    WSB_REQUIRE(kernel_id >= 0 && kernel_id <= WSB_LW4, WSB_ENOTSUP, "kernel id %d has no per-edge channels", kernel_id);
    if (P == 0) return WSB_OK;
    WSB_REQUIRE(pair_q && pair_e && pos && sup && verts && edges && out, WSB_EINVAL, "null buffer");
    hipStream_t s = wsb::as_stream(stream);
    if (form == WSB_FORM_PLAIN) {
        WSB_REQUIRE(plan, WSB_EINVAL, "the plain form needs a plan");
        PlainPairArgs pa{plan_view(plan), pair_q, pair_e, P, nullptr, nullptr, 0, nq, ne, pos, sup, verts, edges, out};
        int dev = 0;
        WSB_HIP(hipGetDevice(&dev));
        const int blocks = std::min((P + 255) / 256, grid_for(dev));
        hipLaunchKernelGGL(plain_pairs_kernel, dim3(blocks), dim3(256), 0, s, pa);
        WSB_HIP(hipGetLastError());
        return WSB_OK;
    }
    PairArgs a{pair_q, pair_e, P, nq, ne, pos, sup, verts, edges, out};
    if (form == WSB_FORM_FP64) return dispatch_pairs<closed::FP64>(kernel_id, a, s);
    if (form == WSB_FORM_MIXED) return dispatch_pairs<closed::MIXED>(kernel_id, a, s);
    return wsb::set_error(WSB_ENOTSUP, "unknown form %d", form);
}
