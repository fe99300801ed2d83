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
enum wsb_kernel {
    WSB_CUBIC = 0, WSB_W2 = 1, WSB_W4 = 2, WSB_W6 = 3,
    /* the wall-operation kernels of the reference's full aggregate layout (scene/provider.py:73-89), 2-D only:
     * the cover vector's cone (1 - q), the wall Laplacian's L = W'/r of w2 / w4, and W^5 of w2 / w4 (tensile; the
     * gradient channels 3, 4 only, the other channels are 0; no plain form, no per-edge channels) */
    WSB_CONE = 4, WSB_LW2 = 5, WSB_LW4 = 6, WSB_W2P5 = 7, WSB_W4P5 = 8
};

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
typedef struct wsb_disks2d wsb_disks2d;
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
/* info[10]: K, nx, ny, nbins, ne, nv, ncell_items, device bytes, local winding (1: rows of cells with edges count
 * the crossings of the segment from their cell's reference point; 0: every such row takes the ray over its y-bin),
 * cells with edges that have no reference point (their rows take the ray) */
int wsb_geom2d_info(const wsb_geom2d* g, int64_t* info);
/* device pointers of the uploaded vertices [nv*2] f64 and edges [ne*2] i32 */
int wsb_geom2d_arrays(const wsb_geom2d* g, const double** verts, const int32_t** edges);

/* ---- particle rows sorted by static cell (device counting sort of the cell keys; no counterpart in the reference,
 * which walks rows in input order): order[n] is a permutation of 0..n-1 with the rows of one cell adjacent (in
 * the order their atomic ranks were taken) and the rows outside the grid last.
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
 * fluid, nrm the unit normal into the fluid, both in the local frame; edge index of the closest edge (the lowest id
 * among equal distances, as the reference's argmin).  The first call on a geometry builds its nearest-edge candidate
 * grid on the device (it allocates and waits for the build); later calls allocate nothing. */
int wsb_signed_distance2d(const wsb_geom2d* g, const double* lpos, int m, double support_max,
                          double* d, double* nrm, int32_t* edge, void* stream);

/* ---- several scheme kernels in one pass (the reference's fused stage 1 evaluates the kernel groups of a layout in one
 * launch family, scene/fused.py:521-564): outs[4] (a host array, indexed by kernel id WSB_CUBIC..WSB_W6; nullptr =
 * not requested) receive what wsb_moments2d gives for each kernel, with the chord primitives of every pair (endpoint
 * radii, logarithm, angle, I_m recurrence) computed once per support radius for all of them.  One thread serves one
 * row (its slots in order, the row sums in registers).  order (nullable): the serving order; nwork (device, nullable,
 * needs order): only the first *nwork positions of order hold rows with slots (wsb_count_order2d writes both), and
 * the indicator terms of every row are first stored by a pass over the rows in input order (contiguous stores).  Any
 * order gives the same results.  Forms WSB_FORM_FP64 and WSB_FORM_MIXED; the other arguments as wsb_moments2d. */
int wsb_moments2d_multi(const wsb_geom2d* g, int form, const int32_t* order, const int32_t* nwork, const int32_t* slots,
                        const int32_t* counts, int K, const double* ind, const double* lpos, const double* lsup, int n,
                        const double* rot4, int flags, double* const* outs, void* stream);

/* ---- the sorted-position layout (no counterpart in the reference, whose arrays stay in input order; for callers that
 * sort the rows by cell every step): with order[n] from wsb_sort2d, wsb_adjacency2d_sorted writes the adjacency of
 * row order[i] at position i -- slots[i*K], counts[i] and prow[i*4] = (x, y, h, indicator) of that row -- so that
 * wsb_moments2d_multi_sorted reads every row's record, count and slots contiguously in cell order; it writes each
 * row's outputs at the row itself (outs[k][order[i]*12]).  The values are those of wsb_adjacency2d and
 * wsb_moments2d_multi (same slots, counts, indicator and twelve channels per row); valid[n] (input order, nullable)
 * and pad as wsb_adjacency2d; forms and flags as wsb_moments2d_multi. */
int wsb_adjacency2d_sorted(const wsb_geom2d* g, const double* lpos, const double* lsup, const uint8_t* valid, int n,
                           const int32_t* order, int32_t* slots, int32_t* counts, double* prow, int pad, void* stream);
int wsb_moments2d_multi_sorted(const wsb_geom2d* g, int form, const int32_t* order, const int32_t* slots,
                               const int32_t* counts, int K, const double* prow, int n, const double* rot4, int flags,
                               double* const* outs, void* stream);

/* ---- rows grouped by slot count, the largest count first (device counting sort; no counterpart in the reference,
 * which evaluates rows in input order): order[n] is a permutation of base[0..n) (of 0..n-1 when base is NULL) along
 * which the slot counts never increase (counts at or above 1023 share one bin), the rows with no slot last; *nwork
 * (device, nullable) receives the number of rows with at least one slot.  Passed as `order` and `nwork` to
 * wsb_moments2d_multi, whose threads each serve one row, it gives every wavefront rows of one slot count; it changes
 * which thread serves a row, never a result.  scratch: wsb_count_order2d_scratch_bytes() bytes (counters, cleared by
 * a kernel: no memset). */
size_t wsb_count_order2d_scratch_bytes(void);
int wsb_count_order2d(const int32_t* counts, const int32_t* base, int n, int K, int32_t* order, int32_t* nwork,
                      void* scratch, size_t scratch_bytes, void* stream);

/* ---- wedge areas and directional extremes of the solid within one support (reference scene/fused.py:157-309
 * `_cone_area_kernel`, `_dir_extreme_kernel`, used by `FusedWall.cone_area` / `dir_extreme`, :741-835), FP64, over
 * the slots of wsb_adjacency2d.  The body's pose enters as (cos, sin) of its rotation: world vectors are taken to
 * the body frame by R^T.  wsb_cone_area2d adds area(solid cap disk(x, h) cap wedge(x, axis, half_angle)) to out[q]
 * (half_angle >= pi: the full disk) and the full-disk area to out[n + q]; axes[n*2] world.
 * wsb_dir_extreme2d max-accumulates into out[q] the largest |cos| between the world vector vecs[q] and a direction
 * from the query to a solid point within h (1 where the vector's line meets the solid, 0 for no solid or w = 0). */
int wsb_cone_area2d(const wsb_geom2d* g, const int32_t* slots, const int32_t* counts, int K, const double* ind,
                    const double* lpos, const double* lsup, int n, const double* axes, double pose_cos, double pose_sin,
                    double half_angle, double* out, void* stream);
int wsb_dir_extreme2d(const wsb_geom2d* g, const int32_t* slots, const int32_t* counts, int K, const double* ind,
                      const double* lpos, const double* lsup, int n, const double* vecs, double pose_cos, double pose_sin,
                      double* out, void* stream);

/* ---- conditioned small / sliver triangle (reference edge/warptriangle.py:136-181 `_triangle_kernel` and `channels`,
 * applied by scene/scene.py:407-408 and scene/fused.py:556-564).  For a closed 3-vertex loop, a row whose support h
 * makes the triangle small (longest edge <= 0.02 h) or a sliver (twice the area <= 0.01 longest edge squared) takes
 * an adaptive area quadrature (tolerance relative to the triangle-scaled channels, at most max_panels panels) in
 * place of the edge-local closed form.  tri[8] (host): v0 (x, y), v1 - v0, v2 - v0 of the loop ordered with its
 * shortest edge first (reference `prepare`), |twice the signed area|, its sign (+1 counter-clockwise).
 * values[n*12], estimate[n*12] (nullable; units of h, device channel order, complete: background x plane vector
 * + sign x the area integral), status[n] (nullable): 0 the closed-form route (values and estimate 0), 1 converged,
 * -1 budget exhausted (values NaN).  valid[n] nullable.  out[n*12] (nullable): the rows with status != 0 are
 * overwritten with the values finalised by `flags` (WSB_OUT_PHYSICAL, WSB_OUT_WORLD with rot4): out must hold
 * this representation's contribution alone (WSB_OUT_ACCUMULATE is rejected). */
int wsb_triangle2d(const double* tri, int background, int kernel_id, const double* lpos, const double* lsup,
                   const uint8_t* valid, int n, double tolerance, int max_panels, double* values, double* estimate,
                   int32_t* status, const double* rot4, int flags, double* out, void* stream);
/* the same rows on the host (host pointers; no `out`) */
int wsb_triangle2d_host(const double* tri, int background, int kernel_id, const double* lpos, const double* lsup,
                        const uint8_t* valid, int n, double tolerance, int max_panels, double* values, double* estimate,
                        int32_t* status);

/* ---- solid disks of one body frame: a fibre bundle (reference DiskArrayRep, scene/scene.py:808-888) or one solid
 * disk (ImplicitRep of a DiskBody, scene/scene.py:576-587).  centres[m*2], radii[m] (host).  Builds the static cell
 * list of the centres (cell = support_max + the largest radius; replaces DiskArrayRep.cells, scene/scene.py:869-888)
 * and uploads it; *K = the most disks in any 3 x 3 block of cells (slot capacity per particle). */
int wsb_disks2d_create(const double* centres, const double* radii, int m, double support_max, wsb_disks2d** out,
                       int* K);
int wsb_disks2d_destroy(wsb_disks2d* g);
/* info[5]: K, nx, ny, m, device bytes */
int wsb_disks2d_info(const wsb_disks2d* g, int64_t* info);
/* slots[n*K]: the disks with |centre - x| < h + R in the 3 x 3 cells around the particle's cell, cell by cell (x outer,
 * y inner) in cell-list order, -1 padded; counts[n] (nullable) the filled slots; valid[n] (nullable) 0 = no disks.
 * Replaces fixed_disk_topology / _disk_slots_kernel (scene/fixedadj.py:199-246); bit-exact with it. */
int wsb_disk_adjacency2d(const wsb_disks2d* g, const double* lpos, const double* lsup, const uint8_t* valid, int n,
                         int32_t* slots, int32_t* counts, void* stream);
/* out[n*12]: the twelve channels (device order, units of h unless WSB_OUT_PHYSICAL) of the solid disks of each row's
 * slots, summed in slot order; each disk is complete (its solid, indicator included), so WSB_OUT_NO_INDICATOR is
 * implied.  FP64 radial Gauss quadrature of the arc-angle form (every kernel id, W^5 included).  Replaces
 * warpdisk.disk_channels_device + FusedWall._add_disks (edge/warpdisk.py:98-199, scene/fused.py:567-575), whose
 * five table channels (lam, m1, g0, g1 along / across the centre direction) it extends by the second moments. */
int wsb_disk_moments2d(const wsb_disks2d* g, int kernel_id, const int32_t* slots, const int32_t* counts, int K,
                       const double* lpos, const double* lsup, int n, const double* rot4, int flags, double* out,
                       void* stream);
/* several of the scheme kernels in one pass (as wsb_moments2d_multi for surfaces): outs[4] (a host array, indexed by
 * kernel id WSB_CUBIC..WSB_W6; nullptr = not requested) receive what wsb_disk_moments2d gives for each kernel, every
 * Gauss node's geometry (alpha, its sine and cosine, the node's radius and weight) computed once for all of them.
 * Without the cubic the kernels share their panels; with it every kernel takes the cubic's panels (also cut at its
 * knot r = h / 2: more nodes, the same accuracy).  The values equal wsb_disk_moments2d's to rounding (the compiler
 * fuses the shared products into FMAs differently).  Flags as wsb_disk_moments2d. */
int wsb_disk_moments2d_multi(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K, const double* lpos,
                             const double* lsup, int n, const double* rot4, int flags, double* const* outs, void* stream);

/* the geometric queries of the solid disks within one support, as wsb_cone_area2d / wsb_dir_extreme2d for a surface
 * (replace _disk_cone_area_kernel / _disk_dir_extreme_kernel, scene/fused.py:312-470): per row the swept areas of its
 * slots' disks (closed form per range of ray angles) and the largest |cos| of the directions to their points within h */
int wsb_disk_cone_area2d(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K, const double* lpos,
                         const double* lsup, int n, const double* axes, double pose_cos, double pose_sin,
                         double half_angle, double* out, void* stream);
int wsb_disk_dir_extreme2d(const wsb_disks2d* g, const int32_t* slots, const int32_t* counts, int K, const double* lpos,
                           const double* lsup, int n, const double* vecs, double pose_cos, double pose_sin, double* out,
                           void* stream);
/* the seven channels of one disk (lam, m1 and g0 along the centre direction, g1 and m2 along / across it; units of h)
 * at centre distances D[n] and radii R[n] in units of h, on the host (validation of the device arithmetic) */
int wsb_disk_channels_host(int kernel_id, const double* D, const double* R, int n, double* out);

/* ---- batched evaluation across many resident geometries (continuous batching of a boundary server; no counterpart
 * in the reference).  wsb_geom2d_desc writes wsb_geom2d_desc_bytes() bytes describing a geometry (grid parameters and
 * device pointers); a caller gathers the descriptors of its resident geometries into one device table.  row_desc[n]
 * names each row's geometry (index into that table; a negative index skips the row: no slots, indicator 0, its
 * output untouched).  slots[n*Kstride] with Kstride >= the K of every geometry in the
 * batch.  Each row's slots, indicator and moments are bitwise those of wsb_adjacency2d / wsb_moments2d (unsorted,
 * local frame) on its own geometry.  nrows (device, nullable; also for the cell sort and the plain moments): when
 * given, only the first min(n, *nrows) rows are served while the launch shapes follow n, so a batch of any size up to
 * n can be replayed from one captured graph. */
size_t wsb_geom2d_desc_bytes(void);
int wsb_geom2d_desc(const wsb_geom2d* g, void* out);
int wsb_batch2d_adjacency(const void* descs, const int32_t* row_desc, const int32_t* order, const double* lpos,
                          const double* lsup, int n, int Kstride, int32_t* slots, int32_t* counts, double* ind,
                          const int32_t* nrows, void* stream);
/* rows grouped by (geometry, cell): order[n] for wsb_batch2d_adjacency (the schedule only, never a result), by a
 * counting sort over the bins of every geometry of the table (wsb_desc2d_cell_bases numbers them, nbins in all; rows
 * outside their grid or without a geometry share one last bin).  Its counters are cleared by a kernel and its launch
 * shapes depend on (n, nbins) only: the call can be captured in a HIP graph.  No reference counterpart (the
 * reference sorts nothing for batches of scenes). */
int wsb_desc2d_cell_bases(void* descs_host, int ndesc, int32_t* nbins);
size_t wsb_batch2d_cellsort_scratch_bytes(int n, int nbins);
int wsb_batch2d_cellsort(const void* descs, const int32_t* row_desc, const double* lpos, int n, int nbins, int32_t* order,
                         void* scratch, size_t scratch_bytes, const int32_t* nrows, void* stream);
int wsb_batch2d_moments(const void* descs, const int32_t* row_desc, int kernel_id, int form, const int32_t* slots,
                        const int32_t* counts, int Kstride, const double* ind, const double* lpos, const double* lsup,
                        int n, int flags, double* out, const int32_t* nrows, void* stream);

/* the plain FP64 form of the batched moments: the plan interpreter of the per-scene plain path, one thread per row
 * summing its pairs in slot order (bitwise the per-scene plain rows); the plan's kernel is the batch's kernel */
int wsb_batch2d_plain_moments(const void* descs, const int32_t* row_desc, const wsb_plan* plan, const int32_t* slots,
                              const int32_t* counts, int Kstride, const double* ind, const double* lpos, const double* lsup,
                              int n, int flags, double* out, const int32_t* nrows, void* stream);

/* rows of device-resident requests into a batch and back: segs is a device array of nseg records of
 * wsb_row_segment_bytes() bytes, each {const double* src; double* dst; int64_t first; int64_t count; int64_t value}
 * (src for a gather, dst for a scatter, value for a fill); rows are `width` doubles wide. */
size_t wsb_row_segment_bytes(void);
int wsb_gather_rows(const void* segs, int nseg, int width, double* batch, void* stream);
int wsb_scatter_rows(const void* segs, int nseg, int width, const double* batch, void* stream);
int wsb_fill_rows_i32(const void* segs, int nseg, int32_t* rows, void* stream);

/* ---- solid-sphere closed form (3-D, FP64; reference curvbound/symbolic.py:166-178 outside the ball,
 * mesh3d/radial.py:102-140 on both sides): out[i] = int_ball W dV for a ball of radius R[i] and a query at signed
 * distance d[i] from its surface (d > 0 outside), units of h; NaN when R <= 0 or d < -R. */
int wsb_sphere_closed_form(int kernel_id, const double* R, const double* d, int n, double* out, void* stream);

/* ================================================================ 3-D: closed oriented triangle surfaces
 * The 22 stored moments of baseline(3)[:22] (reference fluxbridge/contract.py:41-46), in request order
 *     m000 m100 m010 m001 m200 m110 m101 m020 m011 m002  Gx Gy Gz  g1xx g1xy g1xz g1yx g1yy g1yz g1zx g1zy g1zz
 * (moments of y^alpha W, kernel gradients of y^alpha, g1[d][j] = int y_d d_j W), y = source - query. */
typedef struct wsb_mesh3d wsb_mesh3d;

enum wsb_mesh_flags {
    WSB_MESH_HOST_ONLY = 1,   /* build the static tables on the host only (host evaluation entry points)            */
    WSB_MESH_SOUP = 2         /* open triangle set: no closedness check, no winding number (adjacency only)          */
};

/* ---- static geometry of one closed surface (reference mesh3d/reference.py:219-250 ClosedSurface checks;
 * fluxbridge/geometry.py:74-80 Geometry.closed_surface).  verts[nv*3], faces[nf*3] (outward orientation: a cavity's
 * faces point into the cavity).  Builds the face records, a grid of cells of support_max (coarser above 2^24 cells)
 * with each cell's candidate faces (vertex box inflated by support_max), the winding number of every face-free cell
 * and the axis walk of every other cell.  info[10] (nullable): K (largest candidate list), nx, ny, nz, nf, cell items,
 * device bytes, face-free components, cells with faces, longest axis walk (items). */
int wsb_mesh3d_create(const double* verts, int nv, const int32_t* faces, int nf, double support_max, int flags,
                      wsb_mesh3d** out, int64_t* info);
int wsb_mesh3d_destroy(wsb_mesh3d* m);
int wsb_mesh3d_info(const wsb_mesh3d* m, int64_t* info);
/* lo_cell[4]: grid origin (x, y, z) and cell size */
int wsb_mesh3d_grid(const wsb_mesh3d* m, double* lo_cell);

/* ---- face adjacency + winding number (reference fluxbridge/device_search.py:68-195 count / fill / sort with
 * each triangle an element; winding number = the solid angle sum of mesh3d/reference.py:255-258 / 4 pi).
 * pos[n*3], sup[n] (<= support_max); order[n] nullable (wsb_sort3d).  offsets[n+1] (CSR, int64) is always written; faces[capacity] receives the
 * kept faces of each particle in ascending order (entries past `capacity` are dropped: the caller compares
 * offsets[n] with it); chi[n] (nullable) the winding number.  A support above support_max (1 + 1e-12), or NaN, is
 * found on the device: offsets[n] then reads -1 and the lists are not valid (one host read of offsets[n] checks both
 * the capacity and the supports).  scratch: wsb_adjacency3d_scratch_bytes(n) bytes. */
size_t wsb_adjacency3d_scratch_bytes(int n);
int wsb_adjacency3d(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                    int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, void* scratch, size_t scratch_bytes,
                    void* stream);
/* the sorted-position layout (no counterpart in the reference; for callers that sort the particles every step): with
 * order[n] from wsb_sort3d, the adjacency of particle order[i] is written at position i -- offsets (CSR over
 * positions), faces, chi[i] and prow[i*4] = (x, y, z, h) of that particle -- so that wsb_moments3d_sorted reads every
 * position's record, pair range and winding number contiguously; the same face lists and winding numbers per particle,
 * bitwise (the cell-list search only). */
int wsb_adjacency3d_sorted(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                           int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, double* prow, void* scratch,
                           size_t scratch_bytes, void* stream);

/* ---- the same face lists (ascending) and winding numbers with the faces found through a tree of the face boxes
 * (reference fluxbridge/device_search.py:125-212: Morton-ordered LBVH over outward-rounded FP32 boxes, stack
 * traversal, the FP64 element test): per particle a stack walk of the tree, leaves through the cell lists' FP32 filter
 * and element test (so the same set), rows of at most 32 faces sorted in registers, longer rows sorted through an LDS
 * bitmap of face ids.  Meshes of at most 262,144 faces (more: WSB_ENOTSUP).  Arguments as wsb_adjacency3d. */
int wsb_adjacency3d_bvh(const wsb_mesh3d* m, const double* pos, const double* sup, const int32_t* order, int n,
                        int64_t* offsets, int32_t* faces, int64_t capacity, double* chi, void* scratch,
                        size_t scratch_bytes, void* stream);

/* ---- particles grouped by grid cell (device counting sort of the cell keys; no counterpart in the reference):
 * order[n] is a permutation of 0..n-1: first the particles whose cell has candidate faces or faces of its own,
 * grouped by cell, then every other particle in input order.  Passed to wsb_adjacency3d it changes which thread
 * serves a particle, never a result. */
size_t wsb_sort3d_scratch_bytes(const wsb_mesh3d* m, int n);
int wsb_sort3d(const wsb_mesh3d* m, const double* pos, int n, int32_t* order, void* scratch, size_t scratch_bytes,
               void* stream);

/* ---- the 22 stored moments per particle from the adjacency (reference fluxbridge/evaluate.py:158-175 batch +
 * reduce, fluxbridge/wp_func.py:260-351 element in compact mode): out[n*22] = sum over the particle's faces of
 * the compact face terms + chi * whole-ball vector; flags WSB_OUT_PHYSICAL (h^(|alpha| - |derivative|)),
 * WSB_OUT_NO_INDICATOR, WSB_OUT_ACCUMULATE.  form WSB_FORM_FP64 or WSB_FORM_MIXED.  Pairs at or past
 * `capacity` are ignored.  `scratch` (device, wsb_moments3d_scratch_bytes(n) bytes) holds the list of the
 * particles with more than 64 faces, each of which one wavefront serves; no allocation, no host synchronisation. */
size_t wsb_moments3d_scratch_bytes(int n);
int wsb_moments3d(const wsb_mesh3d* m, int kernel_id, int form, const int64_t* offsets, const int32_t* faces,
                  int64_t capacity, const double* chi, const double* pos, const double* sup, int n, int flags,
                  double* out, void* scratch, size_t scratch_bytes, void* stream);
/* the 22 moments from the sorted-position layout of wsb_adjacency3d_sorted: a workgroup serves 128 consecutive positions
 * (particles of neighbouring cells), each particle's row stored at the particle (out[order[i]*22]); the values of
 * wsb_moments3d, bitwise.  Arguments otherwise as wsb_moments3d. */
int wsb_moments3d_sorted(const wsb_mesh3d* m, int kernel_id, int form, const int32_t* order, const double* prow,
                         const int64_t* offsets, const int32_t* faces, int64_t capacity, const double* chi, int n,
                         int flags, double* out, void* scratch, size_t scratch_bytes, void* stream);

/* ---- the six Hessian entries of baseline(3)[22:28] (reference fluxbridge/contract.py:41-46, the request (0, (i, j)),
 * i <= j; mesh3d/reference.py:296-305 kernel_hessian at alpha = 0): out[n*6] = int_V d_i d_j W dV in the order (0,0)
 * (0,1) (0,2) (1,1) (1,2) (2,2), units of h^-2 of the support-unit kernel unless WSB_OUT_PHYSICAL (h^-2 applied),
 * from the face lists of wsb_adjacency3d (no winding term: the whole ball's Hessian is zero).  Flags
 * WSB_OUT_PHYSICAL, WSB_OUT_ACCUMULATE; forms WSB_FORM_FP64 and WSB_FORM_MIXED; scratch as wsb_moments3d. */
int wsb_hessian3d(const wsb_mesh3d* m, int kernel_id, int form, const int64_t* offsets, const int32_t* faces,
                  int64_t capacity, const double* pos, const double* sup, int n, int flags, double* out,
                  void* scratch, size_t scratch_bytes, void* stream);

/* ---- wsb_moments3d without allocation or scratch argument.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The list of the
 * particles with more than 64 faces lives in a file-scope device array of 8 MiB + 256 bytes per GPU (held until the
 * process ends), so a call serves at most wsb_moments3d_static_max_rows() = 2,097,152 particles (more: WSB_EINVAL).
 * Every call on one GPU shares that GPU's array (its counter is cleared by a kernel at the start of each call): calls
 * that overlap on one GPU race.  Arguments and results as wsb_moments3d. */
int64_t wsb_moments3d_static_max_rows(void);
int wsb_moments3d_static(const wsb_mesh3d* m, int kernel_id, int form, const int64_t* offsets, const int32_t* faces,
                         int64_t capacity, const double* chi, const double* pos, const double* sup, int n, int flags,
                         double* out, void* stream);

/* ---- host evaluations of the same arithmetic (validation; every pointer is host memory) */
int wsb_adjacency3d_host(const wsb_mesh3d* m, const double* pos, const double* sup, int n, int64_t* offsets,
                         int32_t* faces, int64_t capacity, double* chi);
int wsb_winding3d_exact_host(const wsb_mesh3d* m, const double* pos, int n, double* chi);
/* out[P*23]: face accumulators (M0, M1[3], V_Phi, M2[6], G[3], g1[9]) of each (query, face) pair, units of h; form
 * WSB_FORM_MIXED evaluates the single-precision angles and square-root seeds of the mixed form on the host */
int wsb_face_terms3d_host(const wsb_mesh3d* m, int kernel_id, int form, const int32_t* pair_face, const double* pair_pos,
                          const double* pair_sup, int P, double* out);
/* out[P*6]: the Hessian face accumulators of each (query, face) pair (before the factor 1/pi and h^-2) */
int wsb_face_hessian3d_host(const wsb_mesh3d* m, int kernel_id, int form, const int32_t* pair_face,
                            const double* pair_pos, const double* pair_sup, int P, double* out);
int wsb_finalize22_host(int kernel_id, const double* acc, const double* chi, const double* sup, int n, int flags,
                        double* out);

#ifdef __cplusplus
}
#endif
#endif
