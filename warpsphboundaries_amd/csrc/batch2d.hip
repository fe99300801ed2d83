// Batched 2-D evaluation across many resident geometries: every row names its geometry by an index into a device
// table of geometry descriptors, so one adjacency launch and one moment launch serve the requests of many scenes
// (continuous batching of a boundary server; no counterpart in the reference, which evaluates one scene per call).
// The per-row arithmetic is the per-scene one: the same slot predicate (operation order kept, no contraction), the
// same winding number, the same closed form and row finalisation, so each row's result is bitwise the result of
// the per-scene entry points (the slot kernel is compiled without contraction, as geom2d.hip; the moment kernel with
// the contraction of moments2d.hip).

#include <cstring>

#include <hipcub/hipcub.hpp>

#include "edge2d_device.h"
#include "geom2d_device.h"
#include "wsb_internal.h"

// device copy of one geometry's grid parameters and array pointers
struct GeoDesc2 {
    double clo0, clo1, cell, y0, dy, bg;
    int32_t nx, ny, nb, K, ne;
    int32_t cbase;              // the first bin of this geometry's cells in a batch cell sort (wsb_desc2d_cell_bases)
    const int32_t* cstart;
    const int32_t* citems;
    const double4* cxy;
    const double* cind;
    const int32_t* cempty;
    const int32_t* bstart;
    const double4* bxy;
    const double* V;
    const int32_t* E;
    const double4* erec;
    const double4* etan;
    const double4* cref;        // nullptr: no local winding (see adjacency2d_kernel)
};

namespace {

constexpr int TILEB = 256;

struct BatchAdjArgs {
    const GeoDesc2* descs;
    const int32_t* row_desc;
    const int32_t* order;       // nullable: thread t serves row order[t] (rows grouped by geometry and cell)
    const double* lpos;
    const double* lsup;
    int n;
    int Kstride;
    int32_t* slots;
    int32_t* counts;
    double* ind;
    const int32_t* nrows;       // nullable (device): only the first min(n, *nrows) rows are served
};

// one thread per row, the fixed-capacity slot fill of the row's own geometry (reference scene/fixedadj.py:67-147)
__global__ __launch_bounds__(256) void batch_adjacency2d_kernel(BatchAdjArgs a) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (a.nrows ? min(a.n, *a.nrows) : a.n)) return;
    const int r = a.order ? a.order[t] : t;
    const int di = a.row_desc[r];
    if (di < 0) {                      // a row without a geometry in this pass
        a.counts[r] = 0;
        a.ind[r] = 0.0;
        return;
    }
    const GeoDesc2& d = a.descs[di];
    const double px = a.lpos[2 * r], py = a.lpos[2 * r + 1];
    const double h = a.lsup[r];
    int n = 0;
    double ind;
    const double fx = floor((px - d.clo0) / d.cell);
    const double fy = floor((py - d.clo1) / d.cell);
    const bool inside = fx >= 0.0 && fx < (double)d.nx && fy >= 0.0 && fy < (double)d.ny;
    if (inside) {
        const int cid = (int)fx * d.ny + (int)fy;
        const double reach = h * (1.0 + 1e-9) + 1e-14 * (1.0 + fabs(px) + fabs(py));
        const bool empty = d.cempty[cid] != 0;
        double4 q = double4{0.0, 0.0, 0.0, 0.0};
        if (!empty && d.cref) q = d.cref[cid];
        const bool local = q.w != 0.0;
        int dw = 0;
        bool unsure = false;
        for (int i = d.cstart[cid]; i < d.cstart[cid + 1]; ++i) {
            const double4 rec = d.cxy[i];
            const double ax = rec.x, ay = rec.y;
            const double bx = rec.z, by = rec.w;
            if (local) local_winding_edge(ax, ay, bx, by, px, py, q.x, q.y, dw, unsure);
            if (fmin(ax, bx) - px > reach || px - fmax(ax, bx) > reach || fmin(ay, by) - py > reach || py - fmax(ay, by) > reach)
                continue;
            const double dx = bx - ax;
            const double dyy = by - ay;
            double t = ((px - ax) * dx + (py - ay) * dyy) / fmax(dx * dx + dyy * dyy, 1.0e-300);
            t = fmin(fmax(t, 0.0), 1.0);
            const double ex = px - (ax + t * dx);
            const double ey = py - (ay + t * dyy);
            if (sqrt(ex * ex + ey * ey) < h) {
                if (n < a.Kstride) a.slots[(int64_t)r * a.Kstride + n] = d.citems[i];
                ++n;
            }
        }
        ind = empty              ? d.cind[cid]
              : local && !unsure ? q.z + (double)dw
                                 : winding_point(px, py, h, d.V, d.E, d.ne, d.bstart, d.bxy, d.y0, d.dy, d.nb, d.bg);
    } else {
        ind = d.bg;
    }
    a.counts[r] = n < a.Kstride ? n : a.Kstride;
    a.ind[r] = ind;
}

struct BatchTileArgs {
    const GeoDesc2* descs;
    const int32_t* row_desc;
    const int32_t* slots;
    const int32_t* counts;
    int Kstride;
    const double* ind;
    const double* lpos;
    const double* lsup;
    int n;
    int flags;
    double* out;
    const int32_t* nrows;       // nullable (device): only the first min(n, *nrows) rows are served
};

// the tile path of the per-scene moments with the row's geometry taken from the descriptor table
template <int KID, int FORM>
__global__ __launch_bounds__(TILEB) void batch_tile12_kernel(BatchTileArgs a) {
    using BlockScan = hipcub::BlockScan<int, TILEB>;
    __shared__ typename BlockScan::TempStorage scan_tmp;
    __shared__ int s_off[TILEB + 1];
    __shared__ double s_px[TILEB], s_py[TILEB], s_h[TILEB], s_ih[TILEB];
    __shared__ double s_res[wsb::NCH][TILEB];
    __shared__ int s_desc[TILEB];
    const int tid = threadIdx.x;
    const int row = blockIdx.x * TILEB + tid;
    const int n = a.nrows ? min(a.n, *a.nrows) : a.n;
    if ((int64_t)blockIdx.x * TILEB >= n) return;          // a whole workgroup past the rows
    const bool live = row < n;
    const int di = live ? a.row_desc[row] : -1;
    const int cnt = di >= 0 ? min(max(a.counts[row], 0), a.Kstride) : 0;
    double h = 1.0;
    if (live) {
        s_px[tid] = a.lpos[2 * row];
        s_py[tid] = a.lpos[2 * row + 1];
        h = a.lsup[row];
        s_h[tid] = h;
        s_ih[tid] = 1.0 / h;
        s_desc[tid] = di;
    }
    int off, total;
    BlockScan(scan_tmp).ExclusiveSum(cnt, off, total);
    s_off[tid] = off;
    if (tid == 0) s_off[TILEB] = total;
    __syncthreads();
    double acc[wsb::NCH];
#pragma unroll
    for (int k = 0; k < wsb::NCH; ++k) acc[k] = 0.0;
    for (int base = 0; base < total; base += TILEB) {
        const int j = base + tid;
        double c[wsb::NCH];
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) c[k] = 0.0;
        if (j < total) {
            int lo = 0, hi = TILEB;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= j) lo = mid;
                else hi = mid;
            }
            const int r = lo;
            const GeoDesc2& d = a.descs[s_desc[r]];
            const int e = a.slots[(int64_t)(blockIdx.x * TILEB + r) * a.Kstride + (j - s_off[r])];
            if (e >= 0 && e < d.ne) {
                const double x = s_px[r], y = s_py[r];
                const double4 ab = d.erec[e];
                const double4 tn = d.etan[e];
                closed::pair_channels<KID, FORM>(ab.x - x, ab.y - y, ab.z - x, ab.w - y, s_h[r], s_ih[r], tn.x, tn.y, tn.z, c);
            }
        }
#pragma unroll
        for (int k = 0; k < wsb::NCH; ++k) s_res[k][tid] = c[k];
        __syncthreads();
        const int b0 = max(off, base), b1 = min(off + cnt, base + TILEB);
        for (int jj = b0; jj < b1; ++jj) {
#pragma unroll
            for (int k = 0; k < wsb::NCH; ++k) acc[k] += s_res[k][jj - base];
        }
        __syncthreads();
    }
    if (di >= 0) finalize_row(acc, a.ind[row], h, plane_of<KID>(), nullptr, a.flags, a.out + (int64_t)row * wsb::NCH);
}

template <int FORM>
int dispatch_batch_tile(int kid, const BatchTileArgs& a, hipStream_t s) {
    const dim3 g((a.n + TILEB - 1) / TILEB), b(TILEB);
    switch (kid) {
        case WSB_CUBIC: hipLaunchKernelGGL((batch_tile12_kernel<WSB_CUBIC, FORM>), g, b, 0, s, a); break;
        case WSB_W2: hipLaunchKernelGGL((batch_tile12_kernel<WSB_W2, FORM>), g, b, 0, s, a); break;
        case WSB_W4: hipLaunchKernelGGL((batch_tile12_kernel<WSB_W4, FORM>), g, b, 0, s, a); break;
        case WSB_W6: hipLaunchKernelGGL((batch_tile12_kernel<WSB_W6, FORM>), g, b, 0, s, a); break;
        default: return wsb::set_error(WSB_ENOTSUP, "unknown kernel id %d", kid);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

// ------------------------------------------------------------------------------------------------ plain FP64 path
// the plan interpreter of the per-scene plain path, one thread per row: the row's pairs are evaluated and summed in
// slot order (the order of the per-scene row sums), so each row is bitwise the per-scene plain result; no pair
// scratch, whatever the largest cell population of the batch's geometries
struct BatchPlainArgs {
    plain::Plan plan;
    const GeoDesc2* descs;
    const int32_t* row_desc;
    const int32_t* slots;
    const int32_t* counts;
    int Kstride;
    const double* ind;
    const double* pos;
    const double* sup;
    int n;
    double plane[12];
    int flags;
    double* out;
    const int32_t* nrows;
};

__global__ __launch_bounds__(256) void batch_plain_rows_kernel(BatchPlainArgs a) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (a.nrows ? min(a.n, *a.nrows) : a.n)) return;
    const int di = a.row_desc[r];
    if (di < 0) return;
    const GeoDesc2& d = a.descs[di];
    double acc[wsb::NCH];
    for (int k = 0; k < wsb::NCH; ++k) acc[k] = 0.0;
    const double h = a.sup[r];
    const double x = a.pos[2 * r], y = a.pos[2 * r + 1];
    const int c = min(max(a.counts[r], 0), a.Kstride);
    for (int t = 0; t < c; ++t) {
        const int e = a.slots[(int64_t)r * a.Kstride + t];
        double ch[wsb::NCH];
        for (int k = 0; k < wsb::NCH; ++k) ch[k] = 0.0;
        if (e >= 0 && e < d.ne) {
            const int v0 = d.E[2 * e], v1 = d.E[2 * e + 1];
            const double px = (d.V[2 * v0] - x) / h, py = (d.V[2 * v0 + 1] - y) / h;
            const double qx = (d.V[2 * v1] - x) / h, qy = (d.V[2 * v1 + 1] - y) / h;
            plain::pair_channels(a.plan, px, py, qx, qy, ch);
        }
        for (int k = 0; k < wsb::NCH; ++k) acc[k] += ch[k];
    }
    finalize_row(acc, a.ind[r], h, a.plane, nullptr, a.flags, a.out + (int64_t)r * wsb::NCH);
}

// ------------------------------------------------------------------------------------------------ request rows
// rows of many device-resident requests gathered into one batch and the batch's rows scattered back: seg[s] =
// (source or destination pointer, first batch row, row count); rows of segment s are seg rows [first, first+count)
struct RowSeg {
    const void* src;
    void* dst;
    int64_t first;
    int64_t count;
    int64_t value;          // the row value of wsb_fill_rows_i32
};

// blockIdx.x = segment, blockIdx.y = one of gridDim.y chunks of it: gather `width` doubles per row (positions 2,
// supports 1) or scatter (outputs 12)
constexpr int SEG_CHUNKS = 64;

__global__ __launch_bounds__(256) void gather_rows_kernel(const RowSeg* seg, int width, double* batch) {
    const RowSeg sg = seg[blockIdx.x];
    const double* src = static_cast<const double*>(sg.src);
    const int64_t m = sg.count * width;
    for (int64_t i = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.y * blockDim.x)
        batch[sg.first * width + i] = src[i];
}

__global__ __launch_bounds__(256) void scatter_rows_kernel(const RowSeg* seg, int width, const double* batch) {
    const RowSeg sg = seg[blockIdx.x];
    double* dst = static_cast<double*>(sg.dst);
    const int64_t m = sg.count * width;
    for (int64_t i = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.y * blockDim.x)
        dst[i] = batch[sg.first * width + i];
}

__global__ __launch_bounds__(256) void fill_rows_kernel(const RowSeg* seg, int32_t* rows) {
    const RowSeg sg = seg[blockIdx.x];
    for (int64_t i = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; i < sg.count; i += (int64_t)gridDim.y * blockDim.x)
        rows[sg.first + i] = (int32_t)sg.value;
}

}  // namespace

extern "C" int wsb_fill_rows_i32(const void* segs, int nseg, int32_t* rows, void* stream) {
    WSB_REQUIRE(nseg >= 0, WSB_EINVAL, "bad fill shape");
    if (nseg == 0) return WSB_OK;
    WSB_REQUIRE(segs && rows, WSB_EINVAL, "null buffer");
    hipLaunchKernelGGL(fill_rows_kernel, dim3(nseg, SEG_CHUNKS), dim3(256), 0, wsb::as_stream(stream),
                       static_cast<const RowSeg*>(segs), rows);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" size_t wsb_geom2d_desc_bytes(void) { return sizeof(GeoDesc2); }

extern "C" int wsb_batch2d_plain_moments(const void* descs, const int32_t* row_desc, const wsb_plan* plan, const int32_t* slots,
                                         const int32_t* counts, int Kstride, const double* ind, const double* lpos,
                                         const double* lsup, int n, int flags, double* out, const int32_t* nrows,
                                         void* stream) {
    WSB_REQUIRE(n >= 0 && Kstride >= 1, WSB_EINVAL, "bad batch shape (n=%d, Kstride=%d)", n, Kstride);
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(descs && row_desc && plan && slots && counts && ind && lpos && lsup && out, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD), WSB_ENOTSUP, "batched moments are evaluated in the local frame");
    BatchPlainArgs a{plain::Plan{plan->radii, plan->cn, plan->cc, plan->E, plan->V, plan->mR, plan->nE, plan->nV},
                     static_cast<const GeoDesc2*>(descs), row_desc, slots, counts, Kstride, ind, lpos, lsup, n, {0}, flags, out,
                     nrows};
    const int kid = plan->kernel_id;
    const double* pl = kid == WSB_CUBIC ? KTab<WSB_CUBIC>::PLANE
                     : kid == WSB_W2    ? KTab<WSB_W2>::PLANE
                     : kid == WSB_W4    ? KTab<WSB_W4>::PLANE
                                        : KTab<WSB_W6>::PLANE;
    for (int k = 0; k < 12; ++k) a.plane[k] = pl[k];
    hipLaunchKernelGGL(batch_plain_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" size_t wsb_row_segment_bytes(void) { return sizeof(RowSeg); }

extern "C" int wsb_gather_rows(const void* segs, int nseg, int width, double* batch, void* stream) {
    WSB_REQUIRE(nseg >= 0 && width >= 1, WSB_EINVAL, "bad gather shape");
    if (nseg == 0) return WSB_OK;
    WSB_REQUIRE(segs && batch, WSB_EINVAL, "null buffer");
    hipLaunchKernelGGL(gather_rows_kernel, dim3(nseg, SEG_CHUNKS), dim3(256), 0, wsb::as_stream(stream),
                       static_cast<const RowSeg*>(segs), width, batch);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_scatter_rows(const void* segs, int nseg, int width, const double* batch, void* stream) {
    WSB_REQUIRE(nseg >= 0 && width >= 1, WSB_EINVAL, "bad scatter shape");
    if (nseg == 0) return WSB_OK;
    WSB_REQUIRE(segs && batch, WSB_EINVAL, "null buffer");
    hipLaunchKernelGGL(scatter_rows_kernel, dim3(nseg, SEG_CHUNKS), dim3(256), 0, wsb::as_stream(stream),
                       static_cast<const RowSeg*>(segs), width, batch);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_geom2d_desc(const wsb_geom2d* g, void* out) {
    WSB_REQUIRE(g && out, WSB_EINVAL, "null argument");
    GeoDesc2 d{};
    d.clo0 = g->clo0;
    d.clo1 = g->clo1;
    d.cell = g->cell;
    d.y0 = g->y0;
    d.dy = g->dy;
    d.bg = (double)g->background;
    d.nx = g->nx;
    d.ny = g->ny;
    d.nb = g->nb;
    d.K = g->K;
    d.ne = g->ne;
    d.cstart = g->cstart;
    d.citems = g->citems;
    d.cxy = g->cxy;
    d.cind = g->cind;
    d.cempty = g->cempty;
    d.cref = g->local ? g->cref : nullptr;
    d.bstart = g->bstart;
    d.bxy = g->bxy;
    d.V = g->verts;
    d.E = g->edges;
    d.erec = g->erec;
    d.etan = g->etan;
    std::memcpy(out, &d, sizeof d);
    return WSB_OK;
}

namespace {
inline size_t al256b(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

// ------------------------------------------------------------------------------------------------ counting sort
// The rows of a batch grouped by (geometry, cell) with a counting sort over the bins of every geometry's grid
// (bin = cbase + cell of the row's own geometry; rows outside their grid or without a geometry share the last bin):
// a kernel clears the counters, one pass counts, a three-level exclusive scan (1,024 bins per workgroup, one
// workgroup over the workgroup totals, one pass adding them back) gives each bin's first position, a last pass places
// every row at its bin's next position.  Fixed launch shapes for a given (n, bins), no memset and no library call:
// the sequence can be captured in a graph.  The order of the rows inside one bin follows the atomics' timing; it only
// decides which thread serves a row, never a result.
namespace {

constexpr int CS_T = 256;
constexpr int CS_PER = 4;                   // bins per thread of a scan workgroup
constexpr int CS_CHUNK = CS_T * CS_PER;     // bins per scan workgroup

__global__ __launch_bounds__(CS_T) void cs_clear_kernel(int32_t* hist, int64_t nb) {
    for (int64_t i = blockIdx.x * (int64_t)CS_T + threadIdx.x; i < nb; i += (int64_t)gridDim.x * CS_T) hist[i] = 0;
}

__global__ __launch_bounds__(CS_T) void cs_count_kernel(const GeoDesc2* descs, const int32_t* row_desc, const double* lpos,
                                                        int n, const int32_t* nrows, int32_t nbins, int32_t* key,
                                                        int32_t* hist) {
    const int r = blockIdx.x * CS_T + threadIdx.x;
    if (nrows) n = min(n, *nrows);
    if (r >= n) return;
    const int di = row_desc[r];
    int32_t k = nbins;
    if (di >= 0) {
        const GeoDesc2& d = descs[di];
        const double fx = floor((lpos[2 * r] - d.clo0) / d.cell);
        const double fy = floor((lpos[2 * r + 1] - d.clo1) / d.cell);
        if (fx >= 0.0 && fx < (double)d.nx && fy >= 0.0 && fy < (double)d.ny) k = d.cbase + (int)fx * d.ny + (int)fy;
    }
    key[r] = k;
    // the last bin (rows outside their grid, often most of a batch) is counted once per wavefront: one atomic
    // per 64 rows instead of one contended atomic per row
    const uint64_t last = __ballot(k == nbins);
    if (k != nbins) atomicAdd(&hist[k], 1);
    else if (__lane_id() == (unsigned)(__ffsll((unsigned long long)last) - 1)) atomicAdd(&hist[nbins], __popcll(last));
}

// exclusive scan of CS_CHUNK bins per workgroup in place; the workgroup's total to tot[blockIdx.x]
__global__ __launch_bounds__(CS_T) void cs_scan_local_kernel(int32_t* hist, int64_t nb, int32_t* tot) {
    __shared__ int32_t s[CS_T];
    const int64_t b0 = blockIdx.x * (int64_t)CS_CHUNK + threadIdx.x * CS_PER;
    int32_t v[CS_PER], sum = 0;
#pragma unroll
    for (int i = 0; i < CS_PER; ++i) {
        v[i] = b0 + i < nb ? hist[b0 + i] : 0;
        sum += v[i];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < CS_T; o <<= 1) {               // inclusive Hillis-Steele over the thread sums
        const int32_t add = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int32_t run = s[threadIdx.x] - sum;
#pragma unroll
    for (int i = 0; i < CS_PER; ++i) {
        if (b0 + i < nb) hist[b0 + i] = run;
        run += v[i];
    }
    if (threadIdx.x == CS_T - 1) tot[blockIdx.x] = s[CS_T - 1];
}

// one workgroup: exclusive scan of the workgroup totals, in chunks of CS_T with a running carry
__global__ __launch_bounds__(CS_T) void cs_scan_top_kernel(int32_t* tot, int nblk) {
    __shared__ int32_t s[CS_T];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nblk; c0 += CS_T) {
        const int i = c0 + threadIdx.x;
        const int32_t v = i < nblk ? tot[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < CS_T; o <<= 1) {
            const int32_t add = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nblk) tot[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += s[CS_T - 1];
        __syncthreads();
    }
}

__global__ __launch_bounds__(CS_T) void cs_scan_add_kernel(int32_t* hist, int64_t nb, const int32_t* tot) {
    const int64_t i = blockIdx.x * (int64_t)CS_T + threadIdx.x;
    if (i < nb) hist[i] += tot[i / CS_CHUNK];
}

__global__ __launch_bounds__(CS_T) void cs_place_kernel(const int32_t* key, int n, const int32_t* nrows, int32_t nbins,
                                                        int32_t* start, int32_t* order) {
    const int r = blockIdx.x * CS_T + threadIdx.x;
    if (nrows) n = min(n, *nrows);
    const int k = r < n ? key[r] : -1;
    // the last bin's rows of a wavefront take one block of positions (one atomic) and their rank inside it
    const uint64_t last = __ballot(k == nbins);
    int first = 0;
    const int leader = last ? __ffsll((unsigned long long)last) - 1 : 0;
    if (last && (int)__lane_id() == leader) first = atomicAdd(&start[nbins], __popcll(last));
    first = __shfl(first, leader);
    if (r >= n) return;
    if (k != nbins) {
        order[atomicAdd(&start[k], 1)] = r;
    } else {
        const uint64_t below = last & ((1ull << __lane_id()) - 1ull);
        order[first + __popcll(below)] = r;
    }
}

struct CSLayout {
    size_t key, hist, tot, total;
    int64_t nb;
    int nblk;
};
CSLayout cs_layout(int n, int32_t nbins) {
    CSLayout L{};
    L.nb = (int64_t)nbins + 1;
    L.nblk = (int)((L.nb + CS_CHUNK - 1) / CS_CHUNK);
    L.key = 0;
    L.hist = al256b(4 * (size_t)n);
    L.tot = al256b(L.hist + 4 * (size_t)L.nb);
    L.total = al256b(L.tot + 4 * (size_t)L.nblk);
    return L;
}

}  // namespace

extern "C" int wsb_desc2d_cell_bases(void* table, int ndesc, int32_t* nbins) {
    WSB_REQUIRE(table && nbins && ndesc >= 0, WSB_EINVAL, "null argument");
    int64_t base = 0;
    for (int i = 0; i < ndesc; ++i) {
        GeoDesc2 d;
        char* p = static_cast<char*>(table) + (size_t)i * sizeof(GeoDesc2);
        std::memcpy(&d, p, sizeof d);
        WSB_REQUIRE(base + (int64_t)d.nx * d.ny < ((int64_t)1 << 30), WSB_EINVAL, "the geometries' cells exceed 2^30 bins");
        d.cbase = (int32_t)base;
        base += (int64_t)d.nx * d.ny;
        std::memcpy(p, &d, sizeof d);
    }
    *nbins = (int32_t)base;
    return WSB_OK;
}

extern "C" size_t wsb_batch2d_cellsort_scratch_bytes(int n, int nbins) {
    if (n <= 0 || nbins < 0) return 0;
    return cs_layout(n, nbins).total;
}

extern "C" int wsb_batch2d_cellsort(const void* descs, const int32_t* row_desc, const double* lpos, int n, int nbins,
                                    int32_t* order, void* scratch, size_t scratch_bytes, const int32_t* nrows, void* stream) {
    WSB_REQUIRE(n >= 0 && nbins >= 0, WSB_EINVAL, "bad batch shape (n=%d, nbins=%d)", n, nbins);
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(descs && row_desc && lpos && order, WSB_EINVAL, "null buffer");
    const CSLayout L = cs_layout(n, nbins);
    WSB_REQUIRE(scratch && scratch_bytes >= L.total, WSB_ENOMEM, "cell sort needs %zu scratch bytes (got %zu)", L.total,
                scratch_bytes);
    char* base = static_cast<char*>(scratch);
    int32_t* key = reinterpret_cast<int32_t*>(base + L.key);
    int32_t* hist = reinterpret_cast<int32_t*>(base + L.hist);
    int32_t* tot = reinterpret_cast<int32_t*>(base + L.tot);
    hipStream_t s = wsb::as_stream(stream);
    const unsigned cb = (unsigned)std::min<int64_t>((L.nb + CS_T - 1) / CS_T, 4096);
    hipLaunchKernelGGL(cs_clear_kernel, dim3(cb), dim3(CS_T), 0, s, hist, L.nb);
    hipLaunchKernelGGL(cs_count_kernel, dim3((n + CS_T - 1) / CS_T), dim3(CS_T), 0, s, static_cast<const GeoDesc2*>(descs),
                       row_desc, lpos, n, nrows, (int32_t)nbins, key, hist);
    hipLaunchKernelGGL(cs_scan_local_kernel, dim3(L.nblk), dim3(CS_T), 0, s, hist, L.nb, tot);
    hipLaunchKernelGGL(cs_scan_top_kernel, dim3(1), dim3(CS_T), 0, s, tot, L.nblk);
    hipLaunchKernelGGL(cs_scan_add_kernel, dim3((unsigned)((L.nb + CS_T - 1) / CS_T)), dim3(CS_T), 0, s, hist, L.nb, tot);
    hipLaunchKernelGGL(cs_place_kernel, dim3((n + CS_T - 1) / CS_T), dim3(CS_T), 0, s, key, n, nrows, (int32_t)nbins, hist,
                       order);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_batch2d_adjacency(const void* descs, const int32_t* row_desc, const int32_t* order, const double* lpos,
                                     const double* lsup, int n, int Kstride, int32_t* slots, int32_t* counts, double* ind,
                                     const int32_t* nrows, void* stream) {
    WSB_REQUIRE(n >= 0 && Kstride >= 1, WSB_EINVAL, "bad batch shape (n=%d, Kstride=%d)", n, Kstride);
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(descs && row_desc && lpos && lsup && slots && counts && ind, WSB_EINVAL, "null buffer");
    BatchAdjArgs a{static_cast<const GeoDesc2*>(descs), row_desc, order, lpos, lsup, n, Kstride, slots, counts, ind, nrows};
    hipLaunchKernelGGL(batch_adjacency2d_kernel, dim3((n + 255) / 256), dim3(256), 0, wsb::as_stream(stream), a);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

extern "C" int wsb_batch2d_moments(const void* descs, const int32_t* row_desc, int kernel_id, int form, const int32_t* slots,
                                   const int32_t* counts, int Kstride, const double* ind, const double* lpos,
                                   const double* lsup, int n, int flags, double* out, const int32_t* nrows, void* stream) {
    WSB_REQUIRE(n >= 0 && Kstride >= 1, WSB_EINVAL, "bad batch shape (n=%d, Kstride=%d)", n, Kstride);
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(descs && row_desc && slots && counts && ind && lpos && lsup && out, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD), WSB_ENOTSUP, "batched moments are evaluated in the local frame");
    BatchTileArgs a{static_cast<const GeoDesc2*>(descs), row_desc, slots, counts, Kstride, ind, lpos, lsup, n, flags, out, nrows};
    hipStream_t s = wsb::as_stream(stream);
    if (form == WSB_FORM_FP64) return dispatch_batch_tile<closed::FP64>(kernel_id, a, s);
    if (form == WSB_FORM_MIXED) return dispatch_batch_tile<closed::MIXED>(kernel_id, a, s);
    return wsb::set_error(WSB_ENOTSUP, "batched moments support the fp64 and mixed forms (form %d)", form);
}
