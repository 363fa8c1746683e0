/*
 * vir_amd -- C ABI of the MI355X (gfx950) projector / SIRT library.
 *
 * Conventions shared by every entry point:
 *   - every array argument is a DEVICE pointer (hipMalloc'd or a torch
 *     tensor's data_ptr()), C-contiguous, float32 / int32 / int64 as typed;
 *   - work is enqueued on `stream` (a hipStream_t, NULL = default stream) and
 *     the call returns without synchronising, unless documented otherwise;
 *   - return 0 on success, VIR_EINVAL for an invalid argument (nothing was
 *     launched), VIR_EHIP for a HIP runtime error; vir_last_error() returns
 *     the calling thread's last message;
 *   - outputs are overwritten unless VIR_ACCUMULATE is passed, in which case
 *     results are added to what the output already holds;
 *   - volumes are (nX, nY, nZ) C-order, voxel (x, y, z) at (x*nY + y)*nZ + z,
 *     the ravel order of the reference's np.ravel_multi_index.
 *
 * Reference interfaces replaced are cited per function as file:line in the
 * vargasp/vir tree.
 */
#ifndef VIR_AMD_H
#define VIR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIR_OK 0
#define VIR_EINVAL (-1)
#define VIR_EHIP (-2)

#define VIR_ACCUMULATE 1u   /* add into the output instead of overwriting  */
#define VIR_SEQUENTIAL 2u   /* plain sequential f32 sums in reference order */
#define VIR_UNGUARDED 4u    /* Siddon FP/BP: vir_siddon_count_guarded gave 0
                               for these rays and grid */
#define VIR_PER_OUTPUT 8u   /* DD 2-D / SF: the one-thread-per-output kernels
                               instead of the tiled ones (same bits) */
#define VIR_SF_UNIT_AMPLITUDE 16u  /* SF parallel modes: footprint amplitude 1
                               instead of 1/(|sin|+|cos|) -- the rectangle
                               footprint of fp_2d_par.py:67-125 */
#define VIR_COMPAT 32u      /* reproduce a reference quirk instead of the
                               intended model: ray-driven FP/BP clamp rays
                               parallel to an axis outside the image onto the
                               border row (rd.py:66-73); SF fan FP keeps a
                               stale o boundary after zero pixels
                               (pd.py:321-325); cone DD FP re-zeroes only part
                               of its u buffer (dd.py:317, 374) */

/* Voxel grid: Grid3d(nPixels, dPixels, origin, float32) of
 * core/core_funcs.py:235-265 plus the snapping epsilon of
 * siddons(..., epsilon) (sys_mat/siddon.py:636). */
typedef struct vir_grid {
    int32_t n[3];
    float d[3];
    float o[3];
    double eps;
} vir_grid;

int vir_version(void);
const char *vir_last_error(void);

/* ---------------------------------------------------------------- Siddon */

/* Pass 1 of siddons(src, trg, ..., flat=True) (sys_mat/siddon.py:636-833):
 * per ray, the number of (voxel, length) entries counts[r] and flags[r]
 * (bit 0: an index fell outside the grid -- the reference raises
 * ValueError in ravel/flat modes, siddon.py:815,821; bit 1: the ray passed
 * the entry/exit gate, siddon.py:759-760).  src, trg: (nrays, 3) float32. */
int vir_siddon_count(const float *src, const float *trg, int64_t nrays,
                     const vir_grid *grid, int32_t *counts, uint8_t *flags,
                     void *stream);

/* Pass 2: write the entries of ray r at [offsets[r], offsets[r+1]).
 * Any of ravel_idx / ix / iy / iz may be NULL.  ravel_idx gets the
 * C-order flat voxel index (int32, as flat_ind of siddon.py:815) of
 * in-grid entries and -1 for out-of-grid ones; ix/iy/iz get the
 * unravelled int32 indices (ravel=False output, siddon.py:823-824). */
int vir_siddon_fill(const float *src, const float *trg, int64_t nrays,
                    const vir_grid *grid, const int64_t *offsets,
                    int32_t *ravel_idx, int32_t *ix, int32_t *iy, int32_t *iz,
                    float *len, void *stream);

/* Adds to *count (one device int32, initialised by the caller) the number of
 * rays whose grid entry/exit division needs the range check, or that meet
 * the grid and need the range-checked plane division or the float64 alpha
 * list (a zero direction component) while tracing, and, on grids whose
 * voxel sizes are powers of two from 1/16 to 16 (walked in voxel units by
 * the unguarded kernels), rays with a coordinate or direction component too
 * small or too large for that walk to round as the world-unit one (near the
 * bounds of the float range only).  When the
 * count is 0, vir_siddon_fp / vir_siddon_bp calls on the same rays and grid
 * may pass VIR_UNGUARDED: their kernels are then built without those paths
 * (same bits, fewer registers).  Passing the flag for other rays gives
 * results that are not the reference's (memory stays in bounds).  No
 * reference counterpart: a precondition check for the flag. */
int vir_siddon_count_guarded(const float *src, const float *trg, int64_t nrays,
                             const vir_grid *grid, int32_t *count, void *stream);

/* On-the-fly forward projection sino[r] = sum_j vol[idx_j] * len_j over the
 * ray's Siddon entries, with no stored matrix.  Same result as
 * siddons(flat=True) followed by the CSR gather of
 * projection/proj_numba.py:11-20 / proj_cuda.py:14-27, accumulated in float64
 * as proj_numba.py:17 does.  The volume has fewer than 2^32 voxels.
 * err_count (may be NULL) receives the number of out-of-grid entries. */
int vir_siddon_fp(const float *src, const float *trg, int64_t nrays,
                  const vir_grid *grid, const float *vol, float *sino,
                  uint32_t flags, int32_t *err_count, void *stream);

/* Exact adjoint of vir_siddon_fp for rays laid out as
 * (nviews, nu, nv), nu*nv < 2^30 -- ray (a, iu, iv) at index (a*nu + iu)*nv + iv --
 * whose view geometry is given by per-view 3x4 projection matrices
 * proj[a*12 .. a*12+11] mapping a world point X to the fractional ray
 * coordinates (iu, iv) = (P0.X~ / P2.X~, P1.X~ / P2.X~) with P2.X~ > 0 in
 * front of the source.  Voxel-brick gather: every brick pulls the rays of its
 * detector footprint, re-traces them with the forward projector's arithmetic
 * and accumulates into on-chip memory in conflict-free colour phases; no
 * atomics, deterministic.  Replaces the scatter backprojectors
 * projection/proj.py:119-150 and projection_c.c:176-219.
 * workspace: device scratch of vir_siddon_bp_workspace() bytes (may be NULL
 * when that size is 0). */
int64_t vir_siddon_bp_workspace(const vir_grid *grid, int32_t nviews);
int vir_siddon_bp(const float *src, const float *trg, int32_t nviews,
                  int32_t nu, int32_t nv, const float *proj,
                  const vir_grid *grid, const float *sino, float *vol,
                  uint32_t flags, void *workspace, void *stream);

/* Per-ray records for vir_siddon_fp_rec / vir_siddon_bp_rec: VIR_SIDDON_RECORD_BYTES per ray
 * (rec 16-byte aligned), the ray's setup as the forward projector computes
 * it -- direction, reciprocals, length, entry/exit alphas, gate and the exact
 * plane slices of siddon.py:705-760.  Built once per ray set and grid.  No
 * reference counterpart: the per-geometry part of the trace, precomputed as
 * the reference precomputes its whole system matrix (siddons(), ~1 KB per
 * ray on the 256^3 cone scan) but at 80 bytes per ray. */
#define VIR_SIDDON_RECORD_BYTES 80
int vir_siddon_ray_records(const float *src, const float *trg, int64_t nrays,
                           const vir_grid *grid, float *rec, void *stream);

/* vir_siddon_fp for rays given by their records (vir_siddon_ray_records of
 * the same rays and grid): same result, bit for bit, without recomputing
 * each ray's setup.  Replaces the same reference interfaces as
 * vir_siddon_fp. */
int vir_siddon_fp_rec(const float *rec, int64_t nrays, const vir_grid *grid,
                      const float *vol, float *sino, uint32_t flags,
                      int32_t *err_count, void *stream);

/* vir_siddon_bp for rays given by their records (same layout, arguments and
 * result, bit for bit): the gather reads each candidate ray's record instead
 * of recomputing its setup.  Replaces the same reference interfaces as
 * vir_siddon_bp. */
int vir_siddon_bp_rec(const float *rec, int32_t nviews, int32_t nu, int32_t nv,
                      const float *proj, const vir_grid *grid,
                      const float *sino, float *vol, uint32_t flags,
                      void *workspace, void *stream);

/* One SIRT iteration (sys_mat/sirt_dd.py:51-72, tests/Test_SIRT/
 * sirt_test.py:25-35) as two projector calls, the elementwise steps folded
 * into the projectors' write-back:
 *   vir_*_fp_res: r = (b - A x) / w_sino where w_sino != 0, else A x;
 *   vir_*_bp_upd: x = x + gamma * u / w_img where w_img != 0 (else
 *                 x + gamma * u), u = A^T r, x updated in place.
 * Same bits as vir_siddon_fp_rec / vir_siddon_bp_rec followed by
 * vir_sirt_residual / vir_sirt_update.  VIR_ACCUMULATE is rejected. */
int vir_siddon_fp_rec_res(const float *rec, int64_t nrays, const vir_grid *grid,
                          const float *x, const float *b, const float *w_sino,
                          float *r, uint32_t flags, void *stream);
int vir_siddon_bp_rec_upd(const float *rec, int32_t nviews, int32_t nu, int32_t nv,
                          const float *proj, const vir_grid *grid, const float *r,
                          const float *w_img, float gamma, float *x, uint32_t flags,
                          void *workspace, void *stream);

/* Batched forms of vir_siddon_fp_rec, vir_siddon_bp_rec and the fused SIRT
 * steps above, for nbatch >= 1 volumes that share the rays: volume j at
 * vol + j*nX*nY*nZ, its sinogram at sino + j*nrays (nrays = nviews*nu*nv
 * for the BP); for the fused steps b, r and x are batched likewise and the
 * weights w_sino, w_img are shared.  They replace the same reference
 * interfaces (projection/proj.py:56-150, sys_mat/sirt_dd.py:51-72) called
 * once per volume.  Each ray is traced once for a group of up to four
 * volumes (FP) or two sinograms (BP), so the trace is paid once per group;
 * every item's result equals the single-volume call on it, bit for bit.
 * err_count receives the rays' out-of-grid entries once.  The BP workspace
 * is vir_siddon_bp_batch_workspace() bytes, or min(nbatch, 2)*nX*nY*nZ*4
 * with VIR_ACCUMULATE on a single-chunk plan. */
int vir_siddon_fp_rec_batch(const float *rec, int64_t nrays, const vir_grid *grid,
                            int32_t nbatch, const float *vol, float *sino,
                            uint32_t flags, int32_t *err_count, void *stream);
int64_t vir_siddon_bp_batch_workspace(const vir_grid *grid, int32_t nviews,
                                      int32_t nbatch);
int vir_siddon_bp_rec_batch(const float *rec, int32_t nviews, int32_t nu, int32_t nv,
                            const float *proj, const vir_grid *grid, int32_t nbatch,
                            const float *sino, float *vol, uint32_t flags,
                            void *workspace, void *stream);
int vir_siddon_fp_rec_res_batch(const float *rec, int64_t nrays, const vir_grid *grid,
                                int32_t nbatch, const float *x, const float *b,
                                const float *w_sino, float *r, uint32_t flags,
                                void *stream);
int vir_siddon_bp_rec_upd_batch(const float *rec, int32_t nviews, int32_t nu, int32_t nv,
                                const float *proj, const vir_grid *grid, int32_t nbatch,
                                const float *r, const float *w_img, float gamma,
                                float *x, uint32_t flags, void *workspace,
                                void *stream);

/* vir_siddon_bp_rec without the workspace argument: the same launches and
 * the same result, bit for bit, staging the view chunks' partial volumes
 * (or, with VIR_ACCUMULATE, the volume) in a fixed device array of
 * VIR_SIDDON_BP_SCRATCH_BYTES owned by the library, one per GPU for the
 * life of the process.  The call allocates nothing (no hipMalloc/hipFree,
 * no allocating library routine).  It returns VIR_EINVAL, launching
 * nothing, when the call needs more staging than that: the staging is
 * vir_siddon_bp_workspace() bytes, or nX*nY*nZ*4 with VIR_ACCUMULATE on a
 * single-chunk plan, and every plan with more than one chunk fits.  Calls
 * that overlap on one GPU (other threads or streams) share the array and
 * race; serialise them.  The array keeps the last call's contents; every
 * byte a call reads was written by that call.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, allocation-free serving loops). */
#define VIR_SIDDON_BP_SCRATCH_BYTES (32u << 20)
int vir_siddon_bp_rec_noalloc(const float *rec, int32_t nviews, int32_t nu, int32_t nv,
                              const float *proj, const vir_grid *grid,
                              const float *sino, float *vol, uint32_t flags,
                              void *stream);

/* ---------------------------------------------- siddons.c (single ray) */

/* Grid of sys_mat/siddons.c: nPixels (VectI) and dPixels (VectF), with the
 * grid lines x_lines (nx+1), y_lines (ny+1), z_lines (nz+1) as its
 * calc_grid_lines writes them (siddons.c:137-148: -n d / 2, then repeated
 * float additions of d), given as device arrays. */
typedef struct vir_sc_grid {
    int32_t nx, ny, nz;
    float dx, dy, dz;
} vir_sc_grid;

/* siddons() of sys_mat/siddons.c:290-341 for nrays rays at once (src, trg:
 * (nrays, 3) float32): counts[r] = the number of VoxelLength rows of ray r
 * (0 when it misses the grid; siddons.c then leaves its row untouched).
 * Same float32 steps as siddons.c, including its range that ignores the z
 * slab (max4/min4, :109-134) and its x-fastest index nx ny z + nx y + x
 * (:253); planes outside [0, N], where siddons.c reads outside its arrays,
 * are not merged. */
int vir_siddons_c_count(const float *src, const float *trg, int64_t nrays,
                        const vir_sc_grid *grid, const float *x_lines,
                        const float *y_lines, const float *z_lines,
                        int32_t *counts, void *stream);

/* The rows themselves: ray r's (idx, length) at [offsets[r], offsets[r+1])
 * (siddons.c:235-261, without the idx = -1 terminator). */
int vir_siddons_c_fill(const float *src, const float *trg, int64_t nrays,
                       const vir_sc_grid *grid, const float *x_lines,
                       const float *y_lines, const float *z_lines,
                       const int64_t *offsets, int32_t *idx, float *len,
                       void *stream);

/* siddons_proj() (siddons.c:345-396, sum_inter :265-287) for nrays rays:
 * ray_sum[r] = sum of vol[idx] * |dST| * dalpha in float32, entry order.
 * vol is (nz, ny, nx) with x fastest.  Entries outside the grid (undefined
 * reads in siddons.c) add nothing; err_count (may be NULL) receives their
 * number. */
int vir_siddons_c_proj(const float *src, const float *trg, int64_t nrays,
                       const vir_sc_grid *grid, const float *x_lines,
                       const float *y_lines, const float *z_lines,
                       const float *vol, float *ray_sum, int32_t *err_count,
                       void *stream);

/* ------------------------------------------------- stored matrix (CSR) */

/* sino[r] = sum_{e in [off[r], off[r+1])} vol[ind[e]] * len[e]
 * (projection_c.c:101-168, proj_numba.py:11-32, proj_cuda.py:14-27).
 * VIR_SEQUENTIAL sums in entry order in plain f32 exactly like
 * forward_proj_ravel_rays_struct; the default is a compensated sum. */
int vir_csr_fp(const int32_t *ind, const float *len, const int64_t *off,
               int64_t nrays, const float *vol, float *sino, uint32_t flags,
               void *stream);

/* Transpose a CSR system matrix into voxel-major (CSC) form: colptr
 * (nvox+1, int64), row (nnz, int32) and clen (nnz, float32), entries of each
 * voxel kept in CSR order (a stable counting sort).  Synchronous. */
int vir_csr_transpose(const int32_t *ind, const float *len, const int64_t *off,
                      int64_t nrays, int64_t nnz, int64_t nvox,
                      int64_t *colptr, int32_t *row, float *clen, void *stream);

/* vol[v] = sum_{e in [colptr[v], colptr[v+1])} sino[row[e]] * clen[e]:
 * the backprojection of projection_c.c:176-219 as an atomic-free gather.
 * VIR_SEQUENTIAL reproduces back_proj_ravel_rays_struct's f32 rounding. */
int vir_csc_bp(const int64_t *colptr, const int32_t *row, const float *clen,
               int64_t nvox, const float *sino, float *vol, uint32_t flags,
               void *stream);

/* ------------------------------------------------------- distance-driven */

/* Geometry of the distance-driven projectors (sys_mat/dd.py).  Volume
 * (nx, ny[, nz]) of cubic voxels d_pix centred at the origin; detector of nu
 * (x nv) bins of pitch du (dv) with offsets su (sv) in bins, boundaries
 * du*(j - nu/2 + su) (dd.py:699, 777, 856-857); na views with per-view
 * cos/sin given as device arrays.  mode: 0 parallel (dd.py:651-725),
 * 1 fan with a flat detector at DSD (dd.py:730-803), 2 fan with an
 * equiangular detector of arc pitch du at radius DSD (no reference
 * counterpart).  The cone projector (dd.py:815-888) ignores mode. */
typedef struct vir_dd_geom {
    int32_t nx, ny, nz;
    int32_t nu, nv, na;
    int32_t mode;
    float d_pix, du, dv, su, sv, DSO, DSD;
} vir_dd_geom;

/* sino[a, u] = out_scale * sum_pixels img * overlap * ray_scale, one thread
 * per bin (gather).  The reference's dd_fp_par_2d / dd_fp_fan_2d use
 * out_scale = d_pix/du (dd.py:725, 803). */
int vir_dd2d_fp(const vir_dd_geom *geom, const float *cosv, const float *sinv,
                const float *img, float *sino, float out_scale, uint32_t flags,
                void *stream);

/* Exact transpose of vir_dd2d_fp (same weights), one thread per pixel.  The
 * reference's dd_bp_par_2d / dd_bp_fan_2d use out_scale = 1/(na*d_pix)
 * (dd.py:721-722, 799-800); out_scale = d_pix/du gives the exact adjoint. */
int vir_dd2d_bp(const vir_dd_geom *geom, const float *cosv, const float *sinv,
                const float *sino, float *img, float out_scale, uint32_t flags,
                void *stream);

/* The fused SIRT steps of vir_siddon_fp_rec_res / vir_siddon_bp_rec_upd on
 * the 2-D distance-driven projector (dd.py:675-803 with sirt_dd.py:51-72). */
int vir_dd2d_fp_res(const vir_dd_geom *geom, const float *cosv, const float *sinv,
                    const float *x, const float *b, const float *w_sino, float *r,
                    float out_scale, uint32_t flags, void *stream);
int vir_dd2d_bp_upd(const vir_dd_geom *geom, const float *cosv, const float *sinv,
                    const float *r, const float *w_img, float gamma, float *x,
                    float out_scale, uint32_t flags, void *stream);

/* Cone-beam DD (dd.py:284-468, 815-888): sino (na, nu, nv); reference
 * scaling out_scale = d_pix^2/(du*dv).  VIR_COMPAT: the reference's sweep
 * with its partially re-zeroed u buffer (dd.py:298-379), one thread per
 * (view, bin) adding in the reference's line/slab order. */
int vir_ddcone_fp(const vir_dd_geom *geom, const float *cosv, const float *sinv,
                  const float *vol, float *sino, float out_scale, uint32_t flags,
                  void *stream);

/* Exact transpose of vir_ddcone_fp (of its default mode; VIR_COMPAT is
 * ignored); reference scaling 1/(na*d_pix). */
int vir_ddcone_bp(const vir_dd_geom *geom, const float *cosv, const float *sinv,
                  const float *sino, float *vol, float out_scale, uint32_t flags,
                  void *stream);

/* --------------------------------------- square (translational) DD */

/* Translational four-sided geometry (sys_mat/dd.py:903-1313): nsrc_p x
 * nsrc_z source positions per side (censpace(nsrc, dsrc, ssrc)), detector
 * boundspace(nu, du, su+ssrc_p) x boundspace(nv, dv, sv+ssrc_z), volume
 * (n, n, nz) of voxel d_pix; sinogram (nsrc_p, nsrc_z, nv, nu, 4). */
typedef struct vir_ddsq_geom {
    int32_t n, nz;
    int32_t nu, nv;
    int32_t nsrc_p, nsrc_z;
    float d_pix, du, dv, dsrc_p, dsrc_z, su, sv, ssrc_p, ssrc_z, DSO, DSD;
} vir_ddsq_geom;

/* sino = out_scale * sum of value * overlap_u * overlap_v over the four
 * sides (dd_fp_square, dd.py:903-1098; the reference uses out_scale =
 * d_pix^2/(du*dv)).  One thread per (source, bin), four sides each. */
int vir_ddsq_fp(const vir_ddsq_geom *geom, const float *img, float *sino,
                float out_scale, uint32_t flags, void *stream);

/* Exact transpose of vir_ddsq_fp's sum (dd_bp_square, dd.py:1104-1313;
 * the reference uses out_scale = 1).  One thread per voxel. */
int vir_ddsq_bp(const vir_ddsq_geom *geom, const float *sino, float *img,
                float out_scale, uint32_t flags, void *stream);

/* ------------------------------------------ separable footprint (pixel-driven) */

/* Geometry of the pixel-driven separable-footprint projectors
 * (sys_mat/pd.py).  mode: 0 parallel 2-D (pd.py:129-234), 1 fan 2-D with a
 * flat detector (pd.py:237-350), 2 cone 3-D (pd.py:353-485), 3 parallel 3-D
 * with detector rows along z (per-slice 2-D parallel footprints times the
 * overlap of each slab with the detector rows; no reference counterpart).
 * Volumes (nbatch, nx, ny[, nz]); sinograms (nbatch, na, nu[, nv]).  ang,
 * cosv, sinv: per-view angle and its float32 cos/sin (device arrays). */
typedef struct vir_sf_geom {
    int32_t nx, ny, nz;
    int32_t nu, nv, na;
    int32_t nbatch;
    int32_t mode;
    float d_pix, du, dv, su, sv, DSO, DSD;
} vir_sf_geom;

/* sino = out_scale * sum over pixels of img * amplitude * overlap/du
 * (* overlap_v/dv).  The reference scales by d_pix (pd.py:234, 350, 485).
 * Parallel 3-D with rows no taller than a voxel runs a tiled kernel that
 * stages each orthogonal line's pixels in LDS (same bits as the one-thread-
 * per-bin gather, which VIR_PER_OUTPUT selects). */
int vir_sf_fp(const vir_sf_geom *geom, const float *ang, const float *cosv,
              const float *sinv, const float *img, float *sino, float out_scale,
              uint32_t flags, void *stream);

/* Exact transpose of vir_sf_fp, one thread per pixel/voxel (parallel 3-D
 * with rows at least half a voxel tall: one wave per pixel column with the
 * footprints of 64 views computed at once, same bits; VIR_PER_OUTPUT selects
 * the per-voxel kernel).  The reference's
 * backprojectors use out_scale = du/(na*d_pix) (2-D, pd.py:231-232, 347-348)
 * and du*dv/(na*d_pix) (cone, pd.py:482-483). */
int vir_sf_bp(const vir_sf_geom *geom, const float *ang, const float *cosv,
              const float *sinv, const float *sino, float *img, float out_scale,
              uint32_t flags, void *stream);

/* -------------------------------- separable-footprint prototypes (sp.py) */

/* Triangle (linear B-spline) footprint fan beam of sys_mat/sp.py:12-87:
 * image (nx, ny) of pixels d_pix centred at the origin, na views with
 * per-view cos/sin (device), nu bins of pitch du with centres
 * du (b - nu/2 + 1/2).  A pixel centre (x, y) projects to u0 = DSD xp /
 * (DSO - yp) with xp = c x + s y, yp = -s x + c y, and weighs bin b by
 * max(0, 1 - |u_b - u0| / h), h = DSD d_pix / (DSO - yp). */
typedef struct vir_sp_geom {
    int32_t nx, ny, nu, na;
    float d_pix, du, DSO, DSD;
} vir_sp_geom;

/* sino (na, nu) = out_scale * sum over pixels of img * weight; the
 * reference's sf_fp_fan_2d uses out_scale = d_pix (sp.py:87).  One thread
 * per bin. */
int vir_sp_fan_fp(const vir_sp_geom *geom, const float *cosv, const float *sinv,
                  const float *img, float *sino, float out_scale, uint32_t flags,
                  void *stream);

/* Exact transpose of vir_sp_fan_fp (same weights), one thread per pixel.
 * No reference counterpart (the reference prototype has no backprojector);
 * the adjoint of sf_fp_fan_2d. */
int vir_sp_fan_bp(const vir_sp_geom *geom, const float *cosv, const float *sinv,
                  const float *sino, float *img, float out_scale, uint32_t flags,
                  void *stream);

/* The separable toy model of sys_mat/sp.py:93-211 (CTGeometry,
 * footprint1D, forward_project, back_project): voxel centres (i - n/2 +
 * 1/2) d along each axis, detector centres (j - n/2 + 1/2) pitch, Ds the
 * source-isocentre and Dd the source-detector distance, f(r) = max(0,
 * 1 - |r|):
 *   proj[v, u] = sum_{x,y,z} vol[x,y,z] f(u/Dd - x/Ds) f(v/Dd - y/Ds) f(z/Ds)
 * (forward_project) and its transpose (back_project), in double.  work:
 * vir_sep_workspace() bytes of device scratch. */
typedef struct vir_sep_geom {
    int32_t nx, ny, nz, nu, nv;
    double dx, dy, dz, su, sv, Ds, Dd;
} vir_sep_geom;

int64_t vir_sep_workspace(const vir_sep_geom *geom);
int vir_sep_fp(const vir_sep_geom *geom, const float *vol, float *proj, void *work,
               void *stream);
int vir_sep_bp(const vir_sep_geom *geom, const float *proj, float *vol, void *work,
               void *stream);

/* ------------------------------------------ filtered backprojection */

/* out[r, i] = sum_j weight[j] sino[r, j] taps[|i - j|] for nrows rows of nu
 * bins (nu <= 4096; weight may be NULL).  With taps[k] = irfft(2 H)[k] of a
 * windowed ramp H this equals filtration.filter_sino (filtration.py:267-329):
 * rFFT of the row zero-padded to z >= 2 nu, times 2 H, inverse, first nu
 * samples.  Not in place. */
int vir_fbp_filter(const float *sino, int64_t nrows, int32_t nu,
                   const float *weight, const float *taps, float *out,
                   void *stream);

/* img = scale * sum over views of the weighted, linearly interpolated
 * filtered projection q (na, nu) at each pixel centre, in the DD projector's
 * geometry (vir_dd_geom; mode 0 parallel: detector coordinate, weight 1;
 * mode 1 flat fan: virtual-detector coordinate DSO p/(DSO - o), weight
 * (DSO/(DSO - o))^2; mode 2 equiangular fan: fan angle, weight 1/L^2).  The
 * pixel-driven backprojection of the reference's fan FBP
 * (processing/reconstruction.py:7-65, processing/proj2.py:580-614). */
int vir_fbp_bp(const vir_dd_geom *geom, const float *cosv, const float *sinv,
               const float *q, float *img, float scale, uint32_t flags,
               void *stream);

/* ------------------------------------------ ray-driven (AW / Joseph) */

/* Geometry of the ray-driven projectors (sys_mat/rd.py).  mode: 0 parallel
 * 2-D (rd.py:472-541), 1 fan 2-D, flat detector (rd.py:566-649), 2 cone 3-D
 * (rd.py:677-790).  joseph: 0 Amanatides-Woo path lengths, 1 Joseph
 * interpolation (0.5-unit samples).  Images (nx, ny[, nz]) centred at the
 * origin with pixel d_pix; sinograms (na, nu[, nv]); cosv/sinv per view
 * (device). */
typedef struct vir_rd_geom {
    int32_t nx, ny, nz;
    int32_t nu, nv, na;
    int32_t mode;
    int32_t joseph;
    float d_pix, du, dv, su, sv, DSO, DSD;
} vir_rd_geom;

/* sino = out_scale * ray sums (AW: value * path length, cone scaled by
 * DSD/sqrt(DSD^2+u^2+v^2); Joseph: 0.5 * interpolated samples); one thread
 * per ray.  The reference's forward projectors use out_scale = 1. */
int vir_rd_fp(const vir_rd_geom *geom, const float *cosv, const float *sinv,
              const float *img, float *sino, float out_scale, uint32_t flags,
              void *stream);

/* The reference's ray-driven backprojection as a gather (one thread per
 * voxel, no atomics): AW path-length splat (rd.py:151-175, 215-247) or
 * Joseph bilinear/trilinear splat (rd.py:356-445).  The reference scales
 * by du/(na*d_pix^2) (2-D, rd.py:538-541, 646-649) or du*dv/na (cone,
 * rd.py:787-788). */
int vir_rd_bp(const vir_rd_geom *geom, const float *cosv, const float *sinv,
              const float *sino, float *img, float out_scale, uint32_t flags,
              void *stream);

/* ------------------------------------------------------ ray geometry */

/* Ray endpoints of a circular orbit, generated on the device
 * (proj_geom.py:61-148 geom_circular).  mode: 0 par, 1 fan, 2 cone.
 * cosv/sinv: per-view cos/sin in double (device); dets_y (nu) and dets_z
 * (nv, or NULL when there is no DetsZ; required for cone) in double.
 * src/trg: float32 (nviews, nu[, nv], 3), the float32 cast that siddons()
 * applies to geom_circular's double endpoints.  Each view's rotation is
 * evaluated as numpy's (N,2)x(2,2) matmul through OpenBLAS dgemm:
 * x' = fma(y, -s, x*c), y' = fma(y, c, x*s). */
int vir_geom_circular(int32_t mode, int64_t nviews, const double *cosv,
                      const double *sinv, int32_t nu, const double *dets_y,
                      int32_t nv, const double *dets_z, double src_iso,
                      double det_iso, float *src, float *trg, void *stream);

/* Ray endpoints of the four-sided square geometry, generated on the device
 * (proj_geom.py:12-57 geom_square): bin b (binsx[b], with neg_binsx[b] =
 * -binsx[b] as the caller's dtype negates it) of side k sits at (b, -d),
 * (-d, b), (-b, d), (d, -b) for d = dist_y and looks along angle a; the
 * target adds the per-angle offsets step_d[a] = cos(a) d 5 and step_e[a] =
 * sin(a) d 5 (double, as numpy evaluates them): side 0 (+D, +E), 1 (+E, +D),
 * 2 (-D, -E), 3 (-E, -D).  src/trg: float32 (nbins, nangs, 4, 3), the
 * float32 cast of geom_square's double endpoints. */
int vir_geom_square(int64_t nbins, const double *binsx, const double *neg_binsx,
                    int64_t nangs, double dist_y, const double *step_d,
                    const double *step_e, float *src, float *trg, void *stream);

/* Synthetic ellipse/ellipsoid phantom on an (nx, ny, nz) grid in
 * FOV-normalised coordinates (the style of phantoms.py): voxel value =
 * sum of rho over the ellipsoids containing its centre.  params: nell x 16
 * floats (centre 3, semi-axes 3, row-major rotation 9, rho), device.  No
 * reference counterpart: the synthetic input of benchmarks and serving. */
int vir_phantom(int32_t nx, int32_t ny, int32_t nz, int32_t nell,
                const float *params, float *out, void *stream);

/* ------------------------------------------------------------- SIRT ops */

/* r[i] = W[i] != 0 ? (b[i] - r[i]) / W[i] : r[i]   (sys_mat/sirt_dd.py:23-26,
 * tests/Test_SIRT/sirt_test.py:29) */
int vir_sirt_residual(const float *b, float *r, const float *W, int64_t n,
                      void *stream);

/* x[i] += gamma * (W[i] != 0 ? u[i] / W[i] : u[i])   (sirt_dd.py:30-33,72;
 * sirt_test.py:33-34) */
int vir_sirt_update(float *x, const float *u, const float *W, float gamma,
                    int64_t n, void *stream);

/* --------------------------------------------------- total variation */

#define VIR_TV_WORK 2048  /* doubles in the work buffer of vir_tv */

/* Smoothed isotropic TV of volumes x (nbatch, nx, ny, nz), z fastest:
 *   value = weight * sum_i (sqrt(|D x|_i^2 + eps^2) - eps)
 * with forward differences D (zero at the last index), and its gradient
 * grad = weight * dTV/dx (overwritten, or added with VIR_ACCUMULATE).
 * grad and value (one device double) may each be NULL; value needs `work`
 * (VIR_TV_WORK device doubles).  The regulariser of the reference's
 * TV-constrained reconstruction (tv.py:200-224 gradient/divergence pair,
 * tests/Test_SIRT/sirt_test.py:37-38) in differentiable form. */
int vir_tv(const float *x, int64_t nbatch, int32_t nx, int32_t ny, int32_t nz,
           float eps, float weight, float *grad, double *value, double *work,
           uint32_t flags, void *stream);

/* Total-variation denoising by Chambolle's projection algorithm (J. Math.
 * Imaging Vis. 20 (2004) 89-97, eq. (9)): u = argmin ||u - g||^2/2 + weight
 * TV(u) (isotropic TV, forward differences zero at the last index) on an
 * image of ndim = 1..3 axes, sizes nx, ny, nz (1 for missing axes), last
 * axis fastest.  Runs `iters` iterations of the dual update on q (ndim
 * planes of the image's size, in/out; start from zeros) with step tau
 * (<= 1/(4 ndim) converges), using v (one image) as scratch; then, when u
 * is not NULL, writes u = g - div q.  dmax (may be NULL): VIR_TV_WORK
 * floats of work; dmax[0] receives the largest |q change| of the last
 * iteration, for a stopping rule.  The denoiser the reference applies to
 * SIRT iterates (tests/Test_SIRT/sirt_test.py:13,38, scikit-image's
 * denoise_tv_chambolle). */
int vir_tv_chambolle(int32_t ndim, int32_t nx, int32_t ny, int32_t nz,
                     const float *g, float *q, float *v, float *u, float weight,
                     float tau, int32_t iters, float *dmax, void *stream);

/* Chambolle-Pock iteration for TV-regularised reconstruction (tv.py:229-294),
 * on volumes (nx, ny, nz) with periodic differences (tv.py:207-224); nz == 1
 * is 2-D.  z holds ndim (2 or 3) planes of the volume's size.
 *   dual TV:   z_c += sigma_z * (roll(ubar,-1,c) - ubar); z /= max(1, |z|)
 *              (the unit-ball projection pyT2, tv.py:24-95, intends)
 *   dual data: q = (q + sigma_q*(Au - g)) / (1 + sigma_q*lam)      (tv.py:283-284)
 *   primal:    u' = u + tau*(div z - ATq); ubar = u' + theta*(u' - u)
 *                                                                   (tv.py:287-291) */
int vir_pdhg_dual_tv(int32_t nx, int32_t ny, int32_t nz, const float *ubar,
                     float *z, float sigma_z, void *stream);
int vir_pdhg_dual_data(int64_t n, const float *Au, const float *g,
                       const float *sigma_q, float lam, float *q, void *stream);
int vir_pdhg_primal(int32_t nx, int32_t ny, int32_t nz, const float *z,
                    const float *ATq, const float *tau, float theta, float *u,
                    float *ubar, void *stream);

/* Least-squares data term: r = a - b (r may be NULL) and
 * value = 0.5 * sum (a - b)^2 (one device double, may be NULL; needs `work`
 * of VIR_TV_WORK doubles), summed in double in a fixed order.  No reference
 * counterpart: the data term of a differentiable reconstruction loss. */
int vir_sqdist(const float *a, const float *b, float *r, int64_t n,
               double *value, double *work, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* VIR_AMD_H */
