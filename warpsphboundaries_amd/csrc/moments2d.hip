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
    const double4* prow;        // non-null: the sorted-position layout (slots, counts, prow by position; out by row)
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

// ------------------------------------------------------------------------------------------------ several kernels
// The scheme kernels cubic, w2, w4, w6 (any subset MASK, bit k = kernel id k) in one pass.  Per pair the chord
// primitives of each support radius -- the clipped chord, the endpoint radii, the logarithm, the angle and the
// differences dI_m -- are computed once (radius 1 for every kernel, radius 1/2 for the cubic's inner block) and
// each kernel's six block sums are accumulated in the same loop over m (rows12_multi_kernel below).
// The FP64 form uses tile12_multi_kernel: a workgroup of MTILE_T threads owns MTILE_R rows, their pairs are packed
// densely over the lanes, each kernel's channels go through LDS to the row sums (slot order, as tile12_kernel); thread
// t keeps the row sums of row t % MTILE_R, channels 3 (t / MTILE_R) .. +2, of every kernel in registers and stages
// them through LDS once at the end, where the factor 1/pi of every channel is applied to the row sum.  (The FP64
// arithmetic needs more registers than one row per thread leaves: the mixed form takes rows12_multi_kernel.)
constexpr int MTILE_T = 256;
constexpr int MTILE_R = 64;

struct BlockSums {
    double AP, VP, AF, VF, BP, BF;
};

template <int MASK>
constexpr int mask_nm() {
    int nm = 0;
    if (MASK & 1) nm = KTab<WSB_CUBIC>::NMAXP > nm ? KTab<WSB_CUBIC>::NMAXP : nm;
    if (MASK & 2) nm = KTab<WSB_W2>::NMAXP > nm ? KTab<WSB_W2>::NMAXP : nm;
    if (MASK & 4) nm = KTab<WSB_W4>::NMAXP > nm ? KTab<WSB_W4>::NMAXP : nm;
    if (MASK & 8) nm = KTab<WSB_W6>::NMAXP > nm ? KTab<WSB_W6>::NMAXP : nm;
    return nm;
}

// the square root of the fused pass: branch-free in the mixed form
template <int FORM>
__device__ __forceinline__ double msqrt(double a) {
    return FORM == closed::MIXED ? closed::xsqrt_mixed_nb(a) : closed::xsqrt<FORM>(a);
}

// one term m of block b of kernel KID.  INTF (the mixed form): the integer form, every sum SC times its value, with
// dIm = dI_m and dIq = dI_m / (m + 2) (m <= NMAXP of the pass) and dRq = (rh^m - rl^m) / m (m >= 2); the coefficients
// PS, FS are integers, so each product takes its coefficient as a literal operand instead of a register (the FP64
// form keeps the profiles P, QP, F, QF with dR = rh^m - rl^m: its workgroup tile kernel has no registers to spare for
// the two extra bases)
template <int KID, int B, bool INTF>
__device__ __forceinline__ void block_term(int m, double dIm, double dIq, double dR, bool have_dI, BlockSums& s) {
    using T = KTab<KID>;
    if (INTF) {
        if (have_dI && m < 16) {
            if (T::PS[B][m] != 0.0) {
                s.AP += T::PS[B][m] * dIm;
                s.VP += T::PS[B][m] * dIq;
            }
            if (T::FS[B][m] != 0.0) {
                s.AF += T::FS[B][m] * dIm;
                s.VF += T::FS[B][m] * dIq;
            }
        }
        if (m >= 2 && m - 2 < 16) {
            if (T::PS[B][m - 2] != 0.0) s.BP += T::PS[B][m - 2] * dR;
            if (T::FS[B][m - 2] != 0.0) s.BF += T::FS[B][m - 2] * dR;
        }
    } else {
        if (have_dI && m < 16) {
            if (T::P[B][m] != 0.0) s.AP += T::P[B][m] * dIm;
            if (T::QP[B][m] != 0.0) s.VP += T::QP[B][m] * dIm;
            if (T::F[B][m] != 0.0) s.AF += T::F[B][m] * dIm;
            if (T::QF[B][m] != 0.0) s.VF += T::QF[B][m] * dIm;
        }
        if (m >= 2 && m - 2 < 16) {
            if (T::QP[B][m - 2] != 0.0) s.BP += T::QP[B][m - 2] * dR;
            if (T::QF[B][m - 2] != 0.0) s.BF += T::QF[B][m - 2] * dR;
        }
    }
}

// the angle weights of block b of kernel KID and the factor of its row sums (1 / (pi SC) in the integer form)
template <int KID, bool INTF>
__device__ __forceinline__ constexpr double mrp(int b) { return INTF ? KTab<KID>::MRPS[b] : KTab<KID>::MRP[b]; }
template <int KID, bool INTF>
__device__ __forceinline__ constexpr double mrf(int b) { return INTF ? KTab<KID>::MRFS[b] : KTab<KID>::MRF[b]; }
template <bool INTF>
__device__ __forceinline__ constexpr double multi_ips(int k) {
    if (!INTF) return WSB_INV_PI;
    return k == 0 ? KTab<WSB_CUBIC>::IPS : k == 1 ? KTab<WSB_W2>::IPS : k == 2 ? KTab<WSB_W4>::IPS : KTab<WSB_W6>::IPS;
}

// the twelve channels of one kernel from its block sums (as closed::pair_channels): mp = sum over blocks of
// MRPS[b] th_b, mf = sum of MRFS[b] th_b (the angle terms, integer form; the cubic's two blocks share the sums' accumulators)
__device__ __forceinline__ void kernel_channels(const BlockSums& s, double mp, double mf, double z, double n0, double n1,
                                                double tx, double ty, double c[wsb::NCH]) {
    const double lam = z * s.VP - mp;
    const double valF = z * s.VF - mf;
    const double zn0 = z * n0, zn1 = z * n1;
    const double YxP = zn0 * s.AP + tx * s.BP, YyP = zn1 * s.AP + ty * s.BP;
    const double YxF = zn0 * s.AF + tx * s.BF, YyF = zn1 * s.AF + ty * s.BF;
    c[0] = lam;
    c[1] = n0 * s.AF;
    c[2] = n1 * s.AF;
    c[3] = -(n0 * s.AP);
    c[4] = -(n1 * s.AP);
    c[5] = lam - n0 * YxP;
    c[6] = -(n1 * YxP);
    c[7] = -(n0 * YyP);
    c[8] = lam - n1 * YyP;
    c[9] = n0 * YxF - valF;
    c[10] = n0 * YyF;
    c[11] = n1 * YyF - valF;
}

// the primitives of one support radius R for one pair; accumulates the sums of the kernels of MASK (SMALL: the
// cubic's inner block only, into the cubic's accumulator S[0]) and sets the chord's angle th.  Branch-free: a lane whose chord is empty (or that has
// no pair, valid false) runs the same instructions on lo = hi = 0, rl = rh = 1, which make every dI_m, every
// difference of powers and the angle exactly zero, so its sums stay zero and the block sums need no merge at a join.
template <int MASK, int FORM, bool SMALL>
__device__ __forceinline__ void radius_sums(double R, bool valid, double z, double az, double zz, double s0, double s1,
                                            double r0sq, double r1sq, double r0, double r1, double& th, BlockSums* S) {
    const bool clo = s0 < 0.0 && r0sq > R * R;
    const bool chi = s1 > 0.0 && r1sq > R * R;
    const double L = msqrt<FORM>(fmax((R - az) * (R + az), 0.0));
    double lo = clo ? -L : s0, hi = chi ? L : s1;
    const bool on = valid && az < R && lo < hi;
    double rl = clo ? R : r0, rh = chi ? R : r1;
    lo = on ? lo : 0.0;
    hi = on ? hi : 0.0;
    rl = on ? rl : 1.0;
    rh = on ? rh : 1.0;
    const double Nh = hi >= 0.0 ? hi + rh : zz, Dh = hi >= 0.0 ? 1.0 : rh - hi;
    const double Nl = lo >= 0.0 ? lo + rl : zz, Dl = lo >= 0.0 ? 1.0 : rl - lo;
    const double lg = log((Nh * Dl) / (Dh * Nl));
    const double dA = (on && z != 0.0) ? lg : 0.0;
    constexpr int NM = SMALL ? KTab<WSB_CUBIC>::NMAXP : mask_nm<MASK>();
    double dIm2 = hi - lo, dIm1 = 0.5 * (hi * rh - lo * rl + zz * dA);   // dI_0, dI_1
    double ph = rh, pl = rl;
#pragma unroll
    for (int m = 0; m <= NM + 2; ++m) {
        if (m >= 2) {
            ph *= rh;
            pl *= rl;
        }
        double dIm = 0.0;
        if (m == 0) dIm = dIm2;
        else if (m == 1) dIm = dIm1;
        else if (m <= NM) {
            dIm = ((hi * ph - lo * pl) + (double)m * zz * dIm2) * (1.0 / (double)(m + 1));
            dIm2 = dIm1;
            dIm1 = dIm;
        }
        constexpr bool INTF = FORM == closed::MIXED;
        const double dR = INTF ? (m >= 2 ? (ph - pl) * RK[m] : 0.0) : ph - pl;
        const double dIq = INTF && m <= NM ? dIm * RK[m + 2] : 0.0;
        const bool have = m <= NM;
        if (SMALL) {
            block_term<WSB_CUBIC, 1, INTF>(m, dIm, dIq, dR, have, S[0]);
        } else {
            if (MASK & 1) block_term<WSB_CUBIC, 0, INTF>(m, dIm, dIq, dR, have && m <= KTab<WSB_CUBIC>::NMAXP, S[0]);
            if (MASK & 2) block_term<WSB_W2, 0, INTF>(m, dIm, dIq, dR, have && m <= KTab<WSB_W2>::NMAXP, S[1]);
            if (MASK & 4) block_term<WSB_W4, 0, INTF>(m, dIm, dIq, dR, have && m <= KTab<WSB_W4>::NMAXP, S[2]);
            if (MASK & 8) block_term<WSB_W6, 0, INTF>(m, dIm, dIq, dR, have && m <= KTab<WSB_W6>::NMAXP, S[3]);
        }
    }
    const double t = FORM == closed::MIXED ? closed::dangle_mixed_nb(z, lo, hi) : closed::dangle_fp64(z, lo, hi);
    th = on ? t : 0.0;
}

struct MultiArgs {
    TileArgs t;                 // t.out unused
    double* out[4];             // per kernel id (nullptr: not requested)
    const int32_t* nwork;       // nullable (device): nwork[0] = the first positions of t.order that are rows with slots
};

template <int MASK, int FORM>
__global__ __launch_bounds__(MTILE_T) __attribute__((amdgpu_waves_per_eu(FORM == closed::MIXED ? 4 : 3))) void tile12_multi_kernel(MultiArgs ma) {
    constexpr bool INTF = FORM == closed::MIXED;
    const TileArgs& a = ma.t;
    using BlockScan = hipcub::BlockScan<int, MTILE_T>;
    __shared__ typename BlockScan::TempStorage scan_tmp;
    __shared__ int s_off[MTILE_T + 1];
    __shared__ int s_row[MTILE_R], s_src[MTILE_R];
    __shared__ double s_px[MTILE_R], s_py[MTILE_R], s_h[MTILE_R], s_ih[MTILE_R];
    __shared__ double s_res[wsb::NCH][MTILE_T];          // one kernel's pair channels; at the end the row sums
    static_assert(4 * wsb::NCH * MTILE_R <= wsb::NCH * MTILE_T && wsb::NCH == 12 && MTILE_T == 4 * MTILE_R,
                  "row sums of four kernels staged in s_res; four groups of three channels");
    const int tid = threadIdx.x;
    const int pos = blockIdx.x * MTILE_R + tid;
    const bool live = tid < MTILE_R && pos < a.n;
    const int row = live ? (a.order ? a.order[pos] : pos) : 0;
    const int src = a.prow ? pos : row;              // where the row's adjacency (and position) is read
    const int cnt = live ? min(max(a.counts[src], 0), a.K) : 0;
    if (live) {
        s_row[tid] = row;
        s_src[tid] = src;
        double h;
        if (a.prow) {
            const double4 pr = a.prow[pos];
            s_px[tid] = pr.x;
            s_py[tid] = pr.y;
            h = pr.z;
        } else {
            s_px[tid] = a.lpos[2 * row];
            s_py[tid] = a.lpos[2 * row + 1];
            h = a.lsup[row];
        }
        s_h[tid] = h;
        s_ih[tid] = 1.0 / h;
    }
    // reduction role: row rr_ of the tile, channels 3 g_ .. 3 g_ + 2
    const int rr_ = tid & (MTILE_R - 1), g_ = tid / MTILE_R;
    double acc[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[k][i] = 0.0;
    int off, total;
    BlockScan(scan_tmp).ExclusiveSum(cnt, off, total);
    s_off[tid] = off;
    if (tid == 0) s_off[MTILE_T] = total;
    __syncthreads();
    for (int base = 0; base < total; base += MTILE_T) {
        const int j = base + tid;
        BlockSums S[4] = {};
        double th1 = 0.0, th05 = 0.0, z = 0.0, n0 = 0.0, n1 = 0.0, tx = 0.0, ty = 0.0;
        // a (64-lane) wavefront whose slots all lie past the tile's pairs skips the pair phase (its sums stay zero)
        if (__builtin_amdgcn_readfirstlane(base + (tid & ~63) < total ? 1 : 0)) {
            // the pair of slot j, branch-free: lanes past the tile's pairs (and invalid slots) take edge 0 and valid =
            // false, which zeroes their sums (the geometry has at least one edge)
            const int jc = min(j, total - 1);
            int lo = 0, hi = MTILE_R;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= jc) lo = mid;
                else hi = mid;
            }
            const int r = lo;
            const int e0 = a.slots[(int64_t)s_src[r] * a.K + (jc - s_off[r])];
            const bool ok = j < total && e0 >= 0 && e0 < a.ne;
            const int e = ok ? e0 : 0;
            const double x = s_px[r], y = s_py[r], h = s_h[r], ih = s_ih[r];
            const double4 ab = a.erec[e];
            const double4 tn = a.etan[e];
            const double ux = ab.x - x, uy = ab.y - y, vx = ab.z - x, vy = ab.w - y;
            tx = tn.x;
            ty = tn.y;
            const double iell = tn.z;
            const bool valid = ok && iell != 0.0;
            double cr = wsb_cross2(ux, uy, vx, vy);
            const double mag = fabs(ux * vy) + fabs(uy * vx);
            if (!(mag > 1e-280 && fabs(cr) > 1.7763568394002505e-15 * mag))
                cr = wsb_cross2(ux / h, uy / h, vx / h, vy / h) * h * h;
            z = cr * ih * iell;
            const double px = ux * ih, py = uy * ih, qx = vx * ih, qy = vy * ih;
            const double s0 = tx * px + ty * py, s1 = tx * qx + ty * qy;
            n0 = ty;
            n1 = -tx;
            const double az = fabs(z), zz = z * z;
            const double r0sq = px * px + py * py, r1sq = qx * qx + qy * qy;
            const double r0 = msqrt<FORM>(r0sq), r1 = msqrt<FORM>(r1sq);
            radius_sums<MASK, FORM, false>(1.0, valid, z, az, zz, s0, s1, r0sq, r1sq, r0, r1, th1, S);
            if (MASK & 1) radius_sums<MASK, FORM, true>(0.5, valid, z, az, zz, s0, s1, r0sq, r1sq, r0, r1, th05, S);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!(MASK & (1 << k))) continue;
            double c[wsb::NCH];
            if (k == 0)
                kernel_channels(S[0], mrp<WSB_CUBIC, INTF>(0) * th1 + mrp<WSB_CUBIC, INTF>(1) * th05,
                                mrf<WSB_CUBIC, INTF>(0) * th1 + mrf<WSB_CUBIC, INTF>(1) * th05, z, n0, n1, tx, ty, c);
            if (k == 1) kernel_channels(S[1], mrp<WSB_W2, INTF>(0) * th1, mrf<WSB_W2, INTF>(0) * th1, z, n0, n1, tx, ty, c);
            if (k == 2) kernel_channels(S[2], mrp<WSB_W4, INTF>(0) * th1, mrf<WSB_W4, INTF>(0) * th1, z, n0, n1, tx, ty, c);
            if (k == 3) kernel_channels(S[3], mrp<WSB_W6, INTF>(0) * th1, mrf<WSB_W6, INTF>(0) * th1, z, n0, n1, tx, ty, c);
#pragma unroll
            for (int q = 0; q < wsb::NCH; ++q) s_res[q][tid] = c[q];
            __syncthreads();
            {
                // slot order; four slots' loads issued before their adds
                const int b0 = max(s_off[rr_], base) - base, b1 = min(s_off[rr_ + 1], base + MTILE_T) - base;
                const double* p0 = s_res[3 * g_];
                const double* p1 = s_res[3 * g_ + 1];
                const double* p2 = s_res[3 * g_ + 2];
                int jj = b0;
                for (; jj + 4 <= b1; jj += 4) {
                    double x0[4], x1[4], x2[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        x0[u] = p0[jj + u];
                        x1[u] = p1[jj + u];
                        x2[u] = p2[jj + u];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        acc[k][0] += x0[u];
                        acc[k][1] += x1[u];
                        acc[k][2] += x2[u];
                    }
                }
                for (; jj < b1; ++jj) {
                    acc[k][0] += p0[jj];
                    acc[k][1] += p1[jj];
                    acc[k][2] += p2[jj];
                }
            }
            __syncthreads();
        }
    }
    // stage the row sums [kernel][channel][row] through s_res, then finalise: thread (kernel k, row r)
    double* stage = &s_res[0][0];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) stage[(k * wsb::NCH + 3 * g_ + i) * MTILE_R + rr_] = acc[k][i];
    __syncthreads();
    const int r = tid & (MTILE_R - 1), k = tid / MTILE_R;
    const int p2 = blockIdx.x * MTILE_R + r;
    if (p2 < a.n && (MASK & (1 << k))) {
        const int rr = s_row[r];
        double acc[wsb::NCH];
#pragma unroll
        for (int q = 0; q < wsb::NCH; ++q) acc[q] = stage[(k * wsb::NCH + q) * MTILE_R + r] * multi_ips<INTF>(k);
        const double* plane = k == 0 ? KTab<WSB_CUBIC>::PLANE : k == 1 ? KTab<WSB_W2>::PLANE : k == 2 ? KTab<WSB_W4>::PLANE
                                                                                                      : KTab<WSB_W6>::PLANE;
        finalize_row(acc, a.prow ? a.prow[s_src[r]].w : a.ind[rr], s_h[r], plane, a.rot4, a.flags,
                     ma.out[k] + (int64_t)rr * wsb::NCH);
    }
}

// ------------------------------------------------------------------------------------------------ one row per lane
// The same closed form with one thread per row: the thread walks its row's slots in order and adds each pair's
// channels of every kernel of MASK to its own row sums in registers (the same sums, in the same slot order, as the
// tile kernel: no LDS, no barrier, no pair-to-row search).  A wavefront costs the largest slot count among its 64
// rows, so the rows are served in the order of wsb_count_order2d (grouped by slot count, largest first): every
// wavefront then has rows of one count, and the rows with no pair (only the indicator term) come last.  The edge
// records of the next slot are loaded while the current pair is evaluated.
constexpr int RROW_T = 256;
constexpr int RHEAVY_T = 256;           // four wavefronts: two workgroups share a compute unit at two waves per SIMD

template <int MASK, int FORM>
__device__ __forceinline__ void serve_row(const MultiArgs& ma, int pos) {
    constexpr bool INTF = FORM == closed::MIXED;
    const TileArgs& a = ma.t;
    const int row = a.order ? a.order[pos] : pos;
    const int64_t src = a.prow ? pos : row;          // where the row's adjacency (and position) is read
    const int cnt = min(max(a.counts[src], 0), a.K);
    double x, y, h;
    if (a.prow) {
        const double4 pr = a.prow[pos];
        x = pr.x;
        y = pr.y;
        h = pr.z;
    } else {
        x = a.lpos[2 * row];
        y = a.lpos[2 * row + 1];
        h = a.lsup[row];
    }
    const double ih = 1.0 / h;
    double acc[4][wsb::NCH];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int q = 0; q < wsb::NCH; ++q) acc[k][q] = 0.0;
    const int32_t* sl = a.slots + src * a.K;
    // two-stage prefetch: the edge records of slot s + 1 and the edge id of slot s + 2 are requested while slot s is
    // evaluated, so a record load never waits on an id load issued in the same iteration
    int e_next = cnt > 0 ? sl[0] : 0;
    bool ok_next = e_next >= 0 && e_next < a.ne;
    double4 ab_next = a.erec[ok_next ? e_next : 0];
    double4 tn_next = a.etan[ok_next ? e_next : 0];
    int e_after = cnt > 1 ? sl[1] : 0;
    for (int s = 0; s < cnt; ++s) {
        const bool ok = ok_next;
        const double4 ab = ab_next, tn = tn_next;
        if (s + 1 < cnt) {
            e_next = e_after;
            ok_next = e_next >= 0 && e_next < a.ne;
            const int en = ok_next ? e_next : 0;
            ab_next = a.erec[en];
            tn_next = a.etan[en];
            if (s + 2 < cnt) e_after = sl[s + 2];
        }
        const double ux = ab.x - x, uy = ab.y - y, vx = ab.z - x, vy = ab.w - y;
        const double tx = tn.x, ty = tn.y, iell = tn.z;
        const bool valid = ok && iell != 0.0;
        double cr = wsb_cross2(ux, uy, vx, vy);
        const double mag = fabs(ux * vy) + fabs(uy * vx);
        if (!(mag > 1e-280 && fabs(cr) > 1.7763568394002505e-15 * mag)) cr = wsb_cross2(ux / h, uy / h, vx / h, vy / h) * h * h;
        const double z = cr * ih * iell;
        const double px = ux * ih, py = uy * ih, qx = vx * ih, qy = vy * ih;
        const double s0 = tx * px + ty * py, s1 = tx * qx + ty * qy;
        const double n0 = ty, n1 = -tx;
        const double az = fabs(z), zz = z * z;
        const double r0sq = px * px + py * py, r1sq = qx * qx + qy * qy;
        const double r0 = msqrt<FORM>(r0sq), r1 = msqrt<FORM>(r1sq);
        BlockSums S[4] = {};
        double th1 = 0.0, th05 = 0.0;
        radius_sums<MASK, FORM, false>(1.0, valid, z, az, zz, s0, s1, r0sq, r1sq, r0, r1, th1, S);
        double c[wsb::NCH];
        // the single-radius kernels first, so that only the cubic's sums live through its second radius
        if (MASK & 2) {
            kernel_channels(S[1], mrp<WSB_W2, INTF>(0) * th1, mrf<WSB_W2, INTF>(0) * th1, z, n0, n1, tx, ty, c);
#pragma unroll
            for (int q = 0; q < wsb::NCH; ++q) acc[1][q] += c[q];
        }
        if (MASK & 4) {
            kernel_channels(S[2], mrp<WSB_W4, INTF>(0) * th1, mrf<WSB_W4, INTF>(0) * th1, z, n0, n1, tx, ty, c);
#pragma unroll
            for (int q = 0; q < wsb::NCH; ++q) acc[2][q] += c[q];
        }
        if (MASK & 8) {
            kernel_channels(S[3], mrp<WSB_W6, INTF>(0) * th1, mrf<WSB_W6, INTF>(0) * th1, z, n0, n1, tx, ty, c);
#pragma unroll
            for (int q = 0; q < wsb::NCH; ++q) acc[3][q] += c[q];
        }
        if (MASK & 1) {
            radius_sums<MASK, FORM, true>(0.5, valid, z, az, zz, s0, s1, r0sq, r1sq, r0, r1, th05, S);
            kernel_channels(S[0], mrp<WSB_CUBIC, INTF>(0) * th1 + mrp<WSB_CUBIC, INTF>(1) * th05,
                            mrf<WSB_CUBIC, INTF>(0) * th1 + mrf<WSB_CUBIC, INTF>(1) * th05, z, n0, n1, tx, ty, c);
#pragma unroll
            for (int q = 0; q < wsb::NCH; ++q) acc[0][q] += c[q];
        }
    }
    const double ind = a.prow ? a.prow[pos].w : a.ind[row];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!(MASK & (1 << k))) continue;
#pragma unroll
        for (int q = 0; q < wsb::NCH; ++q) acc[k][q] *= multi_ips<INTF>(k);
        const double* plane = k == 0 ? KTab<WSB_CUBIC>::PLANE : k == 1 ? KTab<WSB_W2>::PLANE : k == 2 ? KTab<WSB_W4>::PLANE
                                                                                                      : KTab<WSB_W6>::PLANE;
        finalize_row(acc[k], ind, h, plane, a.rot4, a.flags, ma.out[k] + (int64_t)row * wsb::NCH);
    }
}

// one position per thread.  COUNTED (a launch with nwork, its own instantiation so that a profile tells the two
// launches apart): only the first nwork[0] positions are served (the rows with slots of a count order); a workgroup
// past them ends at its first instruction.
template <int MASK, int FORM, bool COUNTED>
__global__ __launch_bounds__(RHEAVY_T) __attribute__((amdgpu_waves_per_eu(2))) void rows12_multi_kernel(MultiArgs ma) {
    const int npos = COUNTED ? min(ma.t.n, ma.nwork[0]) : ma.t.n;
    if ((int64_t)blockIdx.x * RHEAVY_T >= npos) return;
    const int64_t pos = (int64_t)blockIdx.x * RHEAVY_T + threadIdx.x;
    if (pos < npos) serve_row<MASK, FORM>(ma, (int)pos);
}

// with nwork: the indicator term of every row in input order (consecutive threads store consecutive rows); the rows
// with slots are then written again, complete, by rows12_multi_kernel.  Under WSB_OUT_ACCUMULATE only the rows with no
// slot are added here (each row receives exactly one contribution).  A kernel of few registers, so that enough
// wavefronts are resident to keep the stores at HBM speed.
template <int MASK>
__global__ __launch_bounds__(RROW_T) void rows12_indicator_kernel(MultiArgs ma) {
    const TileArgs& a = ma.t;
    __shared__ double stage[RROW_T * wsb::NCH];
    const int64_t r0 = (int64_t)blockIdx.x * RROW_T;
    const int row = (int)r0 + threadIdx.x;
    const int rows = (int)min((int64_t)RROW_T, a.n - r0);
    const bool acc = (a.flags & WSB_OUT_ACCUMULATE) != 0;
    const bool live = row < a.n;
    const double ind = live ? a.ind[row] : 0.0, h = live ? a.lsup[row] : 1.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!(MASK & (1 << k))) continue;
        if (live) {
            double c[wsb::NCH];
#pragma unroll
            for (int q = 0; q < wsb::NCH; ++q) c[q] = 0.0 * WSB_INV_PI;      // the empty row sum, as rows12_multi_kernel
            const double* plane = k == 0 ? KTab<WSB_CUBIC>::PLANE : k == 1 ? KTab<WSB_W2>::PLANE : k == 2 ? KTab<WSB_W4>::PLANE
                                                                                                          : KTab<WSB_W6>::PLANE;
            finalize_row(c, ind, h, plane, a.rot4, a.flags & ~WSB_OUT_ACCUMULATE, stage + threadIdx.x * wsb::NCH);
        }
        __syncthreads();
        double* o = ma.out[k] + r0 * wsb::NCH;
        if (!acc) {
            // the workgroup's rows are one contiguous block of the output: 16-byte stores, consecutive lanes adjacent
            const double2* st2 = reinterpret_cast<const double2*>(stage);
            double2* o2 = reinterpret_cast<double2*>(o);
            for (int j = threadIdx.x; j < rows * wsb::NCH / 2; j += RROW_T) o2[j] = st2[j];
        } else {
            for (int j = threadIdx.x; j < rows * wsb::NCH; j += RROW_T) {
                const int r = (int)r0 + j / wsb::NCH;
                if (min(max(a.counts[r], 0), a.K) == 0) o[j] += stage[j];
            }
        }
        __syncthreads();
    }
}

template <int MASK, int FORM>
int launch_multi(const MultiArgs& a, hipStream_t s) {
    if (FORM == closed::FP64 && !a.nwork) {
        hipLaunchKernelGGL((tile12_multi_kernel<MASK, FORM>), dim3((a.t.n + MTILE_R - 1) / MTILE_R), dim3(MTILE_T), 0, s, a);
        WSB_HIP(hipGetLastError());
        return WSB_OK;
    }
    const dim3 grid((a.t.n + RHEAVY_T - 1) / RHEAVY_T);
    if (a.nwork) {
        hipLaunchKernelGGL((rows12_indicator_kernel<MASK>), dim3((a.t.n + RROW_T - 1) / RROW_T), dim3(RROW_T), 0, s, a);
        hipLaunchKernelGGL((rows12_multi_kernel<MASK, FORM, true>), grid, dim3(RHEAVY_T), 0, s, a);
    } else {
        hipLaunchKernelGGL((rows12_multi_kernel<MASK, FORM, false>), grid, dim3(RHEAVY_T), 0, s, a);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

template <int FORM>
int dispatch_multi(int mask, const MultiArgs& a, hipStream_t s) {
    switch (mask) {
#define WSB_MULTI_CASE(M) \
    case M: return launch_multi<M, FORM>(a, s);
        WSB_MULTI_CASE(1) WSB_MULTI_CASE(2) WSB_MULTI_CASE(3) WSB_MULTI_CASE(4) WSB_MULTI_CASE(5) WSB_MULTI_CASE(6)
        WSB_MULTI_CASE(7) WSB_MULTI_CASE(8) WSB_MULTI_CASE(9) WSB_MULTI_CASE(10) WSB_MULTI_CASE(11) WSB_MULTI_CASE(12)
        WSB_MULTI_CASE(13) WSB_MULTI_CASE(14) WSB_MULTI_CASE(15)
#undef WSB_MULTI_CASE
    }
    return wsb::set_error(WSB_EINVAL, "no kernel requested");
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

extern "C" int wsb_moments2d_multi(const wsb_geom2d* g, int form, const int32_t* order, const int32_t* nwork,
                                   const int32_t* slots, const int32_t* counts, int K, const double* ind, const double* lpos,
                                   const double* lsup, int n, const double* rot4, int flags, double* const* outs,
                                   void* stream) {
    WSB_REQUIRE(g && outs, WSB_EINVAL, "null argument");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes n=%d K=%d", n, K);
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD) || rot4, WSB_EINVAL, "WSB_OUT_WORLD needs rot4");
    WSB_REQUIRE(form == WSB_FORM_FP64 || form == WSB_FORM_MIXED, WSB_ENOTSUP, "several kernels: the fp64 and mixed forms");
    int mask = 0;
    MultiArgs a{};
    for (int k = 0; k < 4; ++k) {
        a.out[k] = outs[k];
        if (outs[k]) mask |= 1 << k;
    }
    WSB_REQUIRE(mask != 0, WSB_EINVAL, "no kernel requested");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(slots && counts && ind && lpos && lsup, WSB_EINVAL, "null buffer");
    WSB_REQUIRE(!nwork || order, WSB_EINVAL, "nwork needs the order it counts (wsb_count_order2d)");
    a.t = TileArgs{order, slots, counts, K, ind, lpos, lsup, n, g->erec, g->etan, g->ne, rot4, flags, nullptr, nullptr};
    a.nwork = nwork;
    hipStream_t s = wsb::as_stream(stream);
    return form == WSB_FORM_FP64 ? dispatch_multi<closed::FP64>(mask, a, s) : dispatch_multi<closed::MIXED>(mask, a, s);
}

extern "C" int wsb_moments2d_multi_sorted(const wsb_geom2d* g, int form, const int32_t* order, const int32_t* slots,
                                          const int32_t* counts, int K, const double* prow, int n, const double* rot4,
                                          int flags, double* const* outs, void* stream) {
    WSB_REQUIRE(g && outs, WSB_EINVAL, "null argument");
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes n=%d K=%d", n, K);
    WSB_REQUIRE(!(flags & WSB_OUT_WORLD) || rot4, WSB_EINVAL, "WSB_OUT_WORLD needs rot4");
    WSB_REQUIRE(form == WSB_FORM_FP64 || form == WSB_FORM_MIXED, WSB_ENOTSUP, "several kernels: the fp64 and mixed forms");
    int mask = 0;
    MultiArgs a{};
    for (int k = 0; k < 4; ++k) {
        a.out[k] = outs[k];
        if (outs[k]) mask |= 1 << k;
    }
    WSB_REQUIRE(mask != 0, WSB_EINVAL, "no kernel requested");
    if (n == 0) return WSB_OK;
    WSB_REQUIRE(order && slots && counts && prow, WSB_EINVAL, "null buffer");
    a.t = TileArgs{order, slots, counts, K, nullptr, nullptr, nullptr, n, g->erec, g->etan, g->ne, rot4, flags, nullptr,
                   reinterpret_cast<const double4*>(prow)};
    a.nwork = nullptr;
    hipStream_t s = wsb::as_stream(stream);
    return form == WSB_FORM_FP64 ? dispatch_multi<closed::FP64>(mask, a, s) : dispatch_multi<closed::MIXED>(mask, a, s);
}

// ------------------------------------------------------------------------------------------------ count order
// Rows grouped by slot count, the largest count first (no counterpart in the reference: it decides which thread serves
// a row, never a result); the rows with no slot come last and their number is left on the device (nwork = the rows
// with slots).  Counting sort over CORD_NB bins (counts at or above CORD_NB - 1 share the last bin), each workgroup
// over a chunk of CORD_CH rows: a histogram (rows with no slot counted per wavefront by a ballot, the others by LDS
// atomics, one global atomic per bin and workgroup), one workgroup's exclusive scan in descending count order, and a
// scatter in which each workgroup reserves one block of positions per bin and places its rows by their ranks.  The
// counters are cleared by a kernel (no memset node when a caller captures the launches).
namespace {

constexpr int CORD_NB = 1024;
constexpr int CORD_T = 256;
constexpr int CORD_CH = 16 * CORD_T;

__device__ __forceinline__ int cord_bin(const int32_t* counts, const int32_t* base, int p, int K) {
    const int r = base ? base[p] : p;
    return min(max(min(counts[r], K), 0), CORD_NB - 1);
}

__global__ __launch_bounds__(CORD_T) void cord_clear_kernel(int32_t* hist, int32_t* cursor) {
    for (int b = threadIdx.x; b < CORD_NB; b += CORD_T) {
        hist[b] = 0;
        cursor[b] = 0;
    }
}

// LDS histogram of the workgroup's chunk
__device__ __forceinline__ void cord_chunk_hist(const int32_t* counts, const int32_t* base, int n, int K, int32_t* lh) {
    for (int b = threadIdx.x; b < CORD_NB; b += CORD_T) lh[b] = 0;
    __syncthreads();
    const int p0 = blockIdx.x * CORD_CH;
    for (int it = 0; it < CORD_CH / CORD_T; ++it) {
        const int p = p0 + it * CORD_T + threadIdx.x;
        const int bin = p < n ? cord_bin(counts, base, p, K) : -1;
        const uint64_t zero = __ballot(bin == 0);
        if ((threadIdx.x & 63) == 0 && zero) atomicAdd(&lh[0], (int)__popcll(zero));
        if (bin > 0) atomicAdd(&lh[bin], 1);
    }
    __syncthreads();
}

__global__ __launch_bounds__(CORD_T) void cord_hist_kernel(const int32_t* counts, const int32_t* base, int n, int K,
                                                           int32_t* hist) {
    __shared__ int32_t lh[CORD_NB];
    cord_chunk_hist(counts, base, n, K, lh);
    for (int b = threadIdx.x; b < CORD_NB; b += CORD_T)
        if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

// off[b] = number of rows in bins above b (descending count order); off[CORD_NB] = rows with at least one slot
__global__ __launch_bounds__(CORD_T) void cord_scan_kernel(const int32_t* hist, int32_t* off, int32_t* nwork) {
    using BlockScan = hipcub::BlockScan<int, CORD_T>;
    __shared__ typename BlockScan::TempStorage tmp;
    constexpr int PER = CORD_NB / CORD_T;
    int v[PER];
    // thread t owns the bins NB-1-(PER t + i): the scan runs from the largest count down
#pragma unroll
    for (int i = 0; i < PER; ++i) v[i] = hist[CORD_NB - 1 - (PER * threadIdx.x + i)];
    int ex[PER];
    BlockScan(tmp).ExclusiveSum(v, ex);
#pragma unroll
    for (int i = 0; i < PER; ++i) off[CORD_NB - 1 - (PER * threadIdx.x + i)] = ex[i];
    if (threadIdx.x == CORD_T - 1) {
        off[CORD_NB] = ex[PER - 1];      // the exclusive sum at bin 0: the rows with at least one slot
        if (nwork) *nwork = ex[PER - 1];
    }
}

__global__ __launch_bounds__(CORD_T) void cord_scatter_kernel(const int32_t* counts, const int32_t* base, int n, int K,
                                                              const int32_t* off, int32_t* cursor, int32_t* order) {
    __shared__ int32_t lh[CORD_NB];
    __shared__ int32_t lbase[CORD_NB];
    cord_chunk_hist(counts, base, n, K, lh);
    for (int b = threadIdx.x; b < CORD_NB; b += CORD_T) {
        lbase[b] = lh[b] ? off[b] + atomicAdd(&cursor[b], lh[b]) : 0;
        lh[b] = 0;
    }
    __syncthreads();
    const int p0 = blockIdx.x * CORD_CH;
    const int lane = threadIdx.x & 63;
    for (int it = 0; it < CORD_CH / CORD_T; ++it) {
        const int p = p0 + it * CORD_T + threadIdx.x;
        const int bin = p < n ? cord_bin(counts, base, p, K) : -1;
        // rows with no slot: the wavefront's block (one LDS atomic) + the set lanes below in the ballot, so they keep
        // their order inside a wavefront
        const uint64_t zero = __ballot(bin == 0);
        int wbase = 0;
        if (lane == 0 && zero) wbase = atomicAdd(&lh[0], (int)__popcll(zero));
        wbase = __shfl(wbase, 0);
        int rank = 0;
        if (bin == 0) rank = wbase + (int)__popcll(zero & ((1ull << lane) - 1ull));
        else if (bin > 0) rank = atomicAdd(&lh[bin], 1);
        if (bin >= 0) order[lbase[bin] + rank] = base ? base[p] : p;
    }
}

}  // namespace

extern "C" size_t wsb_count_order2d_scratch_bytes(void) { return (3 * CORD_NB + 1) * sizeof(int32_t); }

extern "C" int wsb_count_order2d(const int32_t* counts, const int32_t* base, int n, int K, int32_t* order, int32_t* nwork,
                                 void* scratch, size_t scratch_bytes, void* stream) {
    WSB_REQUIRE(n >= 0 && K >= 1, WSB_EINVAL, "bad sizes n=%d K=%d", n, K);
    WSB_REQUIRE(scratch_bytes >= wsb_count_order2d_scratch_bytes(), WSB_EINVAL, "scratch too small: %zu < %zu",
                scratch_bytes, wsb_count_order2d_scratch_bytes());
    WSB_REQUIRE(scratch, WSB_EINVAL, "null scratch");
    int32_t* hist = static_cast<int32_t*>(scratch);
    int32_t* cursor = hist + CORD_NB;
    int32_t* off = cursor + CORD_NB;       // CORD_NB + 1 entries
    hipStream_t s = wsb::as_stream(stream);
    hipLaunchKernelGGL(cord_clear_kernel, dim3(1), dim3(CORD_T), 0, s, hist, cursor);
    if (n > 0) {
        WSB_REQUIRE(counts && order, WSB_EINVAL, "null buffer");
        const int nb = (n + CORD_CH - 1) / CORD_CH;
        hipLaunchKernelGGL(cord_hist_kernel, dim3(nb), dim3(CORD_T), 0, s, counts, base, n, K, hist);
        hipLaunchKernelGGL(cord_scan_kernel, dim3(1), dim3(CORD_T), 0, s, hist, off, nwork);
        hipLaunchKernelGGL(cord_scatter_kernel, dim3(nb), dim3(CORD_T), 0, s, counts, base, n, K, off, cursor, order);
    } else {
        hipLaunchKernelGGL(cord_scan_kernel, dim3(1), dim3(CORD_T), 0, s, hist, off, nwork);
    }
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}
