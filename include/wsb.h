/*
 * warpSPHBoundaries wall-integral hot path on MI355X (gfx950): C ABI.
 *
 * Every function returns an int status (WSB_OK = 0).  On failure the message is available through
 * wsb_last_error (thread-local).  Device buffers are caller-owned raw device pointers (e.g. a torch
 * tensor's data_ptr()); `stream` is a hipStream_t passed as void* (NULL = the default stream).  The
 * evaluate / adjacency entry points allocate nothing and never synchronise the host, so a caller may
 * capture them in a HIP graph; the *_create entry points allocate and upload static data.
 *
 * Units: positions in the local (body) frame, `lsup` = support radius h per particle.  The twelve
 * channels are in the device order
 *     0 lam  1 m10  2 m01  3 Gx  4 Gy  5 g1xx  6 g1xy  7 g1yx  8 g1yy  9 m20  10 m11  11 m02
 * (report order m00 m10 m01 m20 m11 m02 Gx Gy g1xx g1xy g1yx g1yy is the permutation
 * [0,1,2,9,10,11,3,4,5,6,7,8] of it), in units of h unless WSB_OUT_PHYSICAL is set.
 */
#ifndef WSB_H
#define WSB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum wsb_status {
    WSB_OK = 0,
    WSB_EINVAL = 1,      /* invalid argument          -> ValueError           */
    WSB_ENOTSUP = 2,     /* unsupported configuration -> NotImplementedError  */
    WSB_ESTALE = 3,      /* stale adjacency           -> ValueError("stale")  */
    WSB_ENOMEM = 4,      /* allocation / admission    -> MemoryError          */
    WSB_EDEVICE = 5      /* HIP runtime error         -> RuntimeError         */
};

/* kernels: the exact table of the reference (curvbound/kernels.py:41-50) */
enum wsb_kernel { WSB_CUBIC = 0, WSB_W2 = 1, WSB_W4 = 2, WSB_W6 = 3 };

/* arithmetic forms of the evaluation */
enum wsb_form {
    WSB_FORM_PLAIN = 0,  /* all-FP64 plan interpreter, one thread per (particle, edge) pair, pair scratch + row sums */
    WSB_FORM_FP64 = 1,   /* all-FP64 closed form with shared primitive tables, particle tiles, fused row sums    */
    WSB_FORM_MIXED = 2   /* as FP64 with FP32 hardware seeds (sqrt, rcp: one FP64 Newton step) and FP32 angles   */
};

/* output flags of the moment entry points */
enum wsb_out_flags {
    WSB_OUT_ACCUMULATE = 1,   /* out += result (several representations of one body) instead of out = result */
    WSB_OUT_PHYSICAL = 2,     /* scale to physical units: [1,h,h,1/h,1/h,1,1,1,1,h^2,h^2,h^2]                  */
    WSB_OUT_WORLD = 4,        /* rotate vectors by R and tensors by R M R^T (R = rot4 row-major)                */
    WSB_OUT_NO_INDICATOR = 8  /* edge-local sums only (no indicator x plane vector)                             */
};

typedef struct wsb_geom2d wsb_geom2d;
typedef struct wsb_plan wsb_plan;

int wsb_last_error(char* buf, size_t n);
int wsb_version(void);

/* ---- plan (flattened exact twelve-channel term table; reference edge/warpbc.py:41-167, :970-991)
 * radii[nr]; cn[ncoef] powers, cc[ncoef] coefficients; E[nE*7] = (ch, i, a, b, R index, c0, c1);
 * V[nV*4] = (ch, R index, c0, c1), mR[nV].  The plan is uploaded once and used by WSB_FORM_PLAIN. */
int wsb_plan_create(int kernel_id, const double* radii, int nr, const int32_t* cn, const double* cc, int ncoef,
                    const int32_t* E, int nE, const int32_t* V, const double* mR, int nV, wsb_plan** out);
int wsb_plan_destroy(wsb_plan* plan);

/* ---- static 2-D geometry of one surface representation (reference scene/scene.py:210-331):
 * closed loops, counter-clockwise around the solid, background 1 for an exterior solid.  Builds the static
 * edge cell list for support_max (cell = support_max, edge boxes inflated by support_max), the per-cell
 * indicator of edge-free cells and the y-bins of the ray cast, and uploads them.  *K = largest cell
 * population (slot capacity per particle). */
int wsb_geom2d_create(const double* verts, int nv, const int32_t* edges, int ne, int background,
                      double support_max, wsb_geom2d** out, int* K);
int wsb_geom2d_destroy(wsb_geom2d* g);
/* info[8]: K, nx, ny, nbins, ne, nv, ncell_items, device bytes */
int wsb_geom2d_info(const wsb_geom2d* g, int64_t* info);
/* device pointers of the uploaded vertices [nv*2] f64 and edges [ne*2] i32 */
int wsb_geom2d_arrays(const wsb_geom2d* g, const double** verts, const int32_t** edges);

/* ---- particle rows sorted by static cell (device radix sort of the cell keys; no counterpart in the reference,
 * which walks rows in input order): order[n] is a permutation of 0..n-1 with the rows of one cell adjacent.
 * Passing it to the adjacency / moment entry points changes which thread serves a row, never a result. */
size_t wsb_sort2d_scratch_bytes(const wsb_geom2d* g, int n);
int wsb_sort2d(const wsb_geom2d* g, const double* lpos, int n, int32_t* order, void* scratch, size_t scratch_bytes,
               void* stream);

/* ---- fixed-capacity adjacency (reference scene/fixedadj.py:67-147), bit-exact FP64 predicate:
 * slots[n*K]: edges of the particle's cell with point-segment distance < lsup, ascending edge id;
 * counts[n] (nullable); ind[n] winding number + background (0 for invalid rows).  valid[n] nullable (all).
 * order[n] nullable (wsb_sort2d).  pad != 0 writes -1 into the unused slots (the reference layout); pad == 0
 * leaves them untouched. */
int wsb_adjacency2d(const wsb_geom2d* g, const double* lpos, const double* lsup, const uint8_t* valid, int n,
                    const int32_t* order, int32_t* slots, int32_t* counts, double* ind, int pad, void* stream);

/* ---- twelve moments per particle from the adjacency (reference scene/scene.py:384-416 and
 * scene/fused.py:56-154): out[n*12] = sum over the particle's slots (slot order) + ind * plane vector.
 * order[n] (nullable, wsb_sort2d): the tile kernel groups the rows in that order (same results).
 * rot4 (device, row-major R, nullable) for WSB_OUT_WORLD.  WSB_FORM_PLAIN needs `scratch` of
 * wsb_moments2d_scratch_bytes(n, K) bytes (pair list and per-pair channels); the other forms need none. */
size_t wsb_moments2d_scratch_bytes(int n, int K);
int wsb_moments2d(const wsb_geom2d* g, const wsb_plan* plan, int kernel_id, int form, const int32_t* order,
                  const int32_t* slots, const int32_t* counts, int K, const double* ind,
                  const double* lpos, const double* lsup, int n, const double* rot4, int flags,
                  double* out, void* scratch, size_t scratch_bytes, void* stream);

/* ---- edge-local channels of an explicit pair list (reference edge/warpbc.py:919-991, edge_channels +
 * second_channels): out[P*12] in units of h; pairs with pair_e < 0 (or an index outside [0,nq) x [0,ne))
 * give zeros.  pos[nq*2], sup[nq]; verts[*2] and edges[ne*2] are device arrays. */
int wsb_edge_channels2d(const wsb_plan* plan, int kernel_id, int form, const int32_t* pair_q,
                        const int32_t* pair_e, int P, const double* pos, const double* sup, int nq,
                        const double* verts, const int32_t* edges, int ne, double* out, void* stream);

/* ---- signed distance to a surface representation (reference scene/scene.py:1312-1372): d > 0 in the
 * fluid, nrm the unit normal into the fluid, both in the local frame; edge index of the closest edge. */
int wsb_signed_distance2d(const wsb_geom2d* g, const double* lpos, int m, double support_max,
                          double* d, double* nrm, int32_t* edge, void* stream);

/* ---- solid-sphere closed form (3-D, FP64; reference curvbound/symbolic.py:166-178 outside the ball,
 * mesh3d/radial.py:102-140 on both sides): out[i] = int_ball W dV for a ball of radius R[i] and a query at signed
 * distance d[i] from its surface (d > 0 outside), units of h; NaN when R <= 0 or d < -R. */
int wsb_sphere_closed_form(int kernel_id, const double* R, const double* d, int n, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
