// Two-pass counting sort of rows by cell key, shared by the 2-D and 3-D cell sorts.
#pragma once

#include <hipcub/hipcub.hpp>

#include "wsb_internal.h"

namespace {

// Two-pass counting sort by cell without global atomics (grids of at most CS_NC - 1 coarse bins of CS_FMAX cells).
// Coarse bin = cell >> shift, plus one bin for the rows outside the grid.  Pass 1: per chunk of CS_CH rows an LDS
// histogram of the coarse bins, written column by column (bin-major) for one exclusive scan.  Pass 2: per chunk, each
// row takes the next place of its coarse bin from an LDS cursor; a row outside the grid goes straight to its final
// place (the rows outside a 64-row group stay in input order, each group's block placed by one LDS atomic), a row
// inside to a staging list grouped by coarse bin.  Pass 3: one workgroup per coarse bin counts its rows per cell in
// LDS, scans and places them.  Which thread serves a row inside a cell depends on the LDS atomics' timing, never a
// result.
constexpr int CS_T = 256;
constexpr int CS_RPT = 16;
constexpr int CS_CH = CS_T * CS_RPT;
constexpr int CS_NC = 1024;
constexpr int CS_FMAX = 4096;

struct CoarseGrid {
    int shift, ncb;             // ncb coarse bins inside the grid; the rows outside it take bin ncb
};

inline CoarseGrid coarse_grid(int64_t nbins) {
    int sh = 0;
    while ((nbins + (1 << sh) - 1) >> sh > CS_NC - 1) ++sh;
    return {sh, (int)((nbins + (1 << sh) - 1) >> sh)};
}

inline bool two_pass_sort(int64_t nbins) { return (1 << coarse_grid(nbins).shift) <= CS_FMAX; }

__device__ __forceinline__ int lane_rank(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}


// pass 1 from keys computed beforehand (key < 0: no cell; such a row is placed after the others)
__global__ __launch_bounds__(CS_T) void cs_hist_keys_kernel(const int32_t* key, int n, int shift, int nc, int nblk,
                                                            int32_t* ghist) {
    __shared__ int h[CS_NC];
    for (int i = threadIdx.x; i < nc; i += CS_T) h[i] = 0;
    const int64_t c0 = (int64_t)blockIdx.x * CS_CH;
    int k[CS_RPT];
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        k[j] = r < n ? key[r] : -1;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const bool live = c0 + j * CS_T + threadIdx.x < n;
        const uint64_t out = __ballot(live && k[j] < 0);
        if (live && k[j] >= 0) atomicAdd(&h[k[j] >> shift], 1);
        else if (live && lane_rank(out) == 0) atomicAdd(&h[nc - 1], __popcll(out));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += CS_T) ghist[(int64_t)i * nblk + blockIdx.x] = h[i];
}

__global__ __launch_bounds__(CS_T) void cs_scatter_kernel(const int32_t* key, int n, int shift, int nc, int nblk,
                                                          const int32_t* goff, int32_t* stage, int32_t* skey,
                                                          int32_t* order) {
    __shared__ int cur[CS_NC];
    for (int i = threadIdx.x; i < nc; i += CS_T) cur[i] = goff[(int64_t)i * nblk + blockIdx.x];
    const int64_t c0 = (int64_t)blockIdx.x * CS_CH;
    const int lane = threadIdx.x & 63;
    int k[CS_RPT];
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        k[j] = r < n ? key[r] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CS_RPT; ++j) {
        const int64_t r = c0 + j * CS_T + threadIdx.x;
        const bool live = r < n;
        const uint64_t out = __ballot(live && k[j] < 0);
        if (out) {
            const int leader = __builtin_ctzll(out);
            int base = 0;
            if (lane == leader) base = atomicAdd(&cur[nc - 1], __popcll(out));
            base = __shfl(base, leader, 64);
            if (live && k[j] < 0) order[base + lane_rank(out)] = (int)r;
        }
        if (live && k[j] >= 0) {
            const int pos = atomicAdd(&cur[k[j] >> shift], 1);
            stage[pos] = (int)r;
            skey[pos] = k[j];
        }
    }
}

__global__ __launch_bounds__(CS_T) void cs_fine_kernel(const int32_t* stage, const int32_t* skey, const int32_t* goff, int nblk,
                                                       int shift, int32_t* order) {
    __shared__ int fh[CS_FMAX];
    __shared__ int part[CS_T];
    const int cb = blockIdx.x;
    const int s0 = goff[(int64_t)cb * nblk], s1 = goff[(int64_t)(cb + 1) * nblk];
    if (s0 == s1) return;
    const int F = 1 << shift, mask = F - 1;
    for (int i = threadIdx.x; i < F; i += CS_T) fh[i] = 0;
    __syncthreads();
    for (int i0 = s0 + threadIdx.x; i0 < s1; i0 += 4 * CS_T) {
        int f[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) f[u] = i0 + u * CS_T < s1 ? skey[i0 + u * CS_T] & mask : -1;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (f[u] >= 0) atomicAdd(&fh[f[u]], 1);
    }
    __syncthreads();
    // exclusive scan of the F counts: each thread a run of consecutive bins, then the runs' sums
    const int per = (F + CS_T - 1) / CS_T;
    const int b0 = min(F, (int)threadIdx.x * per), b1 = min(F, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; ++b) sum += fh[b];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < CS_T; d <<= 1) {
        const int v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = s0 + part[threadIdx.x] - sum;
    for (int b = b0; b < b1; ++b) {
        const int c = fh[b];
        fh[b] = run;
        run += c;
    }
    __syncthreads();
    for (int i0 = s0 + threadIdx.x; i0 < s1; i0 += 4 * CS_T) {
        int f[4], row[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = i0 + u * CS_T < s1;
            f[u] = in ? skey[i0 + u * CS_T] & mask : -1;
            row[u] = in ? stage[i0 + u * CS_T] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (f[u] >= 0) order[atomicAdd(&fh[f[u]], 1)] = row[u];
    }
}

// scratch of the two-pass sort: key[n] | stage[n] | skey[n] | ghist | goff | scan temporary
struct CsLayout {
    size_t key, stage, skey, ghist, goff, tmp, tmp_bytes, total;
};

inline size_t cs_al256(size_t x) { return (x + 255) & ~(size_t)255; }

inline CsLayout cs_layout(int n, int64_t nbins) {
    CsLayout L{};
    const int nc = coarse_grid(nbins).ncb + 1, nblk = (n + CS_CH - 1) / CS_CH;
    const size_t nh = (size_t)nc * (size_t)nblk;
    size_t tmp = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (const int32_t*)nullptr, (int32_t*)nullptr, (int)nh);
    L.key = 0;
    L.stage = cs_al256(L.key + sizeof(int32_t) * (size_t)n);
    L.skey = cs_al256(L.stage + sizeof(int32_t) * (size_t)n);
    L.ghist = cs_al256(L.skey + sizeof(int32_t) * (size_t)n);
    L.goff = cs_al256(L.ghist + sizeof(int32_t) * nh);
    L.tmp = cs_al256(L.goff + sizeof(int32_t) * nh);
    L.tmp_bytes = tmp;
    L.total = cs_al256(L.tmp + tmp);
    return L;
}

// passes 2 and 3 after pass 1 has filled ghist: the scan, the chunk scatter and the per-coarse-bin placement
inline int cs_place(const CsLayout& L, char* base, int n, int64_t nbins, int32_t* order, hipStream_t s) {
    const CoarseGrid cg = coarse_grid(nbins);
    const int nc = cg.ncb + 1, nblk = (n + CS_CH - 1) / CS_CH;
    int32_t* key = reinterpret_cast<int32_t*>(base + L.key);
    int32_t* stage = reinterpret_cast<int32_t*>(base + L.stage);
    int32_t* skey = reinterpret_cast<int32_t*>(base + L.skey);
    int32_t* ghist = reinterpret_cast<int32_t*>(base + L.ghist);
    int32_t* goff = reinterpret_cast<int32_t*>(base + L.goff);
    size_t tmp = L.tmp_bytes;
    WSB_HIP(hipcub::DeviceScan::ExclusiveSum(base + L.tmp, tmp, ghist, goff, nc * nblk, s));
    hipLaunchKernelGGL(cs_scatter_kernel, dim3(nblk), dim3(CS_T), 0, s, key, n, cg.shift, nc, nblk, goff, stage, skey, order);
    hipLaunchKernelGGL(cs_fine_kernel, dim3(cg.ncb), dim3(CS_T), 0, s, stage, skey, goff, nblk, cg.shift, order);
    WSB_HIP(hipGetLastError());
    return WSB_OK;
}

}  // namespace
