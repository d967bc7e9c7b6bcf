/*
 * nefct_hip.h -- C ABI of the MI355X (gfx950) cone-beam projector library.
 *
 * Drop-in boundary for the TensorFlow custom ops of bill52547/CT-NEF:
 *   ProjectFlat / ProjectCyli / BackProjectFlat / BackProjectCyli
 *     (nefct/cc/siddon_3d/siddon.cc:11-80, launchers siddon.cu.cc:392-448,834-890)
 *   Deform (cpp/deform/deform.cc:12-22, deform.cu.cc:112-126)
 *   FDK weighted backprojection (nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:257-298,
 *     weights/filter of mat_codes/filtering.m and ParkerWeight.m)
 *
 * Conventions (all entry points):
 *   - every array argument is a device pointer owned by the caller, except
 *     grid/center/size, which are HOST arrays of 3 (the reference copies them
 *     device->host synchronously on every call, siddon.cu.cc:403-405);
 *   - volumes are x fastest: image[ix + nx*(iy + ny*iz)];
 *     sinograms are pv[ia + na*(ib + nb*iv)] (shape {nv, nb, na});
 *   - geometry uses the reference siddon-op names: SID = source-to-isocentre,
 *     SAD = source-to-detector (siddon.cu.cc:344-348);
 *     a = (ai + ia) da - na da / 2, b = (bi + ib) db - nb db / 2;
 *   - 'stream' is a hipStream_t passed as void* (NULL = legacy default stream);
 *     no entry point synchronises the host;
 *   - outputs are fully written (zero-filled where the reference zero-fills);
 *   - return 0 (NEFCT_OK) or a negative nefct_status; nothing is launched on
 *     error.  Calls that overlap on one stream-free workspace race.
 */
#ifndef NEFCT_HIP_H
#define NEFCT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NEFCT_OK = 0,
    NEFCT_ERR_NULL = -1,        /* a required pointer is NULL */
    NEFCT_ERR_SHAPE = -2,       /* non-positive or oversized dimension */
    NEFCT_ERR_GEOMETRY = -3,    /* non-finite or degenerate geometry */
    NEFCT_ERR_WORKSPACE = -4,   /* workspace missing or too small */
    NEFCT_ERR_UNSUPPORTED = -5, /* flag combination not supported */
    NEFCT_ERR_LAUNCH = -6,      /* HIP launch failed */
    NEFCT_ERR_SCRATCH = -7      /* beyond the fixed scratch of an allocation-free entry point */
} nefct_status;

/* Flags of the Siddon entry points.  0 = exact Siddon (the adjoint pair).
 * The reference op's behaviour is reproduced by:
 *   forward  NEFCT_REF_FORWARD  = INV_L2 | REF_SLAB
 *   adjoint  NEFCT_REF_ADJOINT  = REF_SLAB | SHORT_SKIP | TIE_GT          */
#define NEFCT_F_INV_L2 1u     /* Q1: weights / L^2, L in x-voxel units (siddon.cu.cc:86) */
#define NEFCT_F_REF_SLAB 2u   /* Q2: reference slab walk and attribution (siddon.cu.cc:41-176) */
#define NEFCT_F_SHORT_SKIP 4u /* Q3: adjoint skips rays with in-plane length < 10 voxels (:472) */
#define NEFCT_F_TIE_GT 8u     /* Q6: adjoint x-walk only if |xd| > |yd| (:480) */
/* Adjoint algorithm selection (ignored by the forward).  All variants apply
 * the same weights; they differ in work mapping and summation.
 *   default                   volume blocks own their voxels in LDS (no global
 *                             atomics); contributions are summed as 64-bit
 *                             fixed-point integers with a power-of-two scale
 *                             from max |projection| (rounding ~1e-9 of the
 *                             largest contribution): bitwise reproducible
 *   NEFCT_F_ADJ_FLOAT_ATOMIC  the same blocks summing with LDS float atomics
 *   NEFCT_F_DETERMINISTIC     with FLOAT_ATOMIC: one wavefront per 16x16x8 block
 *                             (fixed order); without: no change (already so)
 *   NEFCT_F_ADJ_WINDOW        ray tiles with LDS slab windows flushed by global
 *                             float atomics
 * A non-finite projection value makes the fixed-point result NaN. */
#define NEFCT_F_DETERMINISTIC 16u
#define NEFCT_F_ADJ_WINDOW 32u
#define NEFCT_F_ADJ_FLOAT_ATOMIC 0x1000u
/* Block shape of the volume-tile adjoint (x by y by z voxels; 4 wavefronts,
 * 16 for the 16K-voxel blocks, whose fixed-point LDS array is 133 KB) and the
 * spacing of a wavefront's rays; the autotuner picks among these. */
#define NEFCT_F_TILE_MASK 0x700u
#define NEFCT_F_TILE_32x32x16 0x000u  /* default (16 wavefronts) */
#define NEFCT_F_TILE_32x16x8 0x100u
#define NEFCT_F_TILE_32x32x8 0x200u
#define NEFCT_F_TILE_64x32x8 0x300u
#define NEFCT_F_TILE_16x16x32 0x400u
#define NEFCT_F_TILE_64x16x16 0x500u
#define NEFCT_F_LANE_STRIDE2 0x800u  /* neighbouring lanes take every other detector row */
/* Forward work mapping (ignored by the adjoint; same result): rays per
 * workgroup (columns x rows, 256 threads; a wavefront is the first 64), and
 * the number of views an XCD walks together (2^k views for k = 1..6 in the
 * NEFCT_F_FWD_VG field; 0 = chosen from the view count, about 16 degrees of
 * a full scan).  The autotuner picks among these. */
#define NEFCT_F_FWD_TILE_MASK 0x6000u
#define NEFCT_F_FWD_TILE_16x16 0x0000u  /* default */
#define NEFCT_F_FWD_TILE_32x8 0x2000u
#define NEFCT_F_FWD_TILE_8x32 0x4000u
#define NEFCT_F_FWD_TILE_64x4 0x6000u
#define NEFCT_F_FWD_VG_SHIFT 15
#define NEFCT_F_FWD_VG_MASK (7u << NEFCT_F_FWD_VG_SHIFT)
#define NEFCT_REF_FORWARD (NEFCT_F_INV_L2 | NEFCT_F_REF_SLAB)
#define NEFCT_REF_ADJOINT (NEFCT_F_REF_SLAB | NEFCT_F_SHORT_SKIP | NEFCT_F_TIE_GT)

/* Message for a status code (static string). */
const char* nefct_status_string(int status);

/* Shape limits of the Siddon entry points (NEFCT_ERR_SHAPE beyond them):
 * grid sides <= 16384, nv <= 65535, nb <= 262140, na * nb < 2^31 cells per view,
 * padded volume copies below 2^31 bytes.  The FDK backprojection takes
 * na, nb < 2^23 and na * nb < 2^29; the FDK filter na <= 2048. */

/* Bytes of device workspace the Siddon entry points need for this volume and
 * view count (padded volume copies + per-view table).  Replaces the
 * reference's implicit TF allocations; no reference counterpart. */
size_t nefct_siddon_workspace_size(const int grid[3], int nv);

/* Forward projection, flat detector.
 * Replaces ProjectFlatOp::Compute (siddon.cc:129-157) + project_flat
 * (siddon.cu.cc:392-419).  Output projection_value {nv, nb, na}, fully written.
 * NOTE: the reference op passes da for db (siddon.cc:155); this entry point
 * uses db as given (the Python drop-in reproduces the reference call). */
int nefct_project_flat(const float* image, const int grid[3], const float center[3], const float size[3],
                       const float* angles, int nv, float SID, float SAD, int na, float da, float ai,
                       int nb, float db, float bi, float* projection_value, unsigned flags,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Forward projection, cylindrical detector (a in radians).
 * Replaces ProjectCyliOp::Compute (siddon.cc:185-213) + project_cyli (siddon.cu.cc:421-448). */
int nefct_project_cyli(const float* image, const int grid[3], const float center[3], const float size[3],
                       const float* angles, int nv, float SID, float SAD, int na, float da, float ai,
                       int nb, float db, float bi, float* projection_value, unsigned flags,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Adjoint (ray-driven atomic scatter), flat detector.  Output image
 * x fastest, zero-filled then accumulated.
 * Replaces BackProjectFlatOp::Compute (siddon.cc:244-272) + backproject_flat
 * (siddon.cu.cc:834-861). */
int nefct_backproject_flat(const float* projection_value, const int grid[3], const float center[3],
                           const float size[3], const float* angles, int nv, float SID, float SAD,
                           int na, float da, float ai, int nb, float db, float bi, float* image,
                           unsigned flags, void* workspace, size_t workspace_bytes, void* stream);

/* Allocation-free variant of nefct_backproject_flat (no workspace argument).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Calls no hipMalloc/hipFree; its scratch (the
 * fixed-point scale word and the per-view table) is a fixed 128 KB device
 * array per GPU, held until the process ends.  Limits: nv <= 8192 views
 * (NEFCT_ERR_SCRATCH beyond), and NEFCT_F_ADJ_WINDOW is refused
 * (NEFCT_ERR_UNSUPPORTED: it needs the padded-copy workspace); every other
 * flag and result is as nefct_backproject_flat.  Calls that overlap on one
 * GPU share the scratch and race; calls on one stream are ordered. */
int nefct_backproject_flat_noalloc(const float* projection_value, const int grid[3], const float center[3],
                                   const float size[3], const float* angles, int nv, float SID, float SAD,
                                   int na, float da, float ai, int nb, float db, float bi, float* image,
                                   unsigned flags, void* stream);

/* Adjoint, cylindrical detector.  Replaces BackProjectCyliOp::Compute
 * (siddon.cc:306-334) + backproject_cyli (siddon.cu.cc:863-890); reads the
 * sinogram with the flat indexing (the reference's transposed index, Q5 at
 * siddon.cu.cc:830, is not reproduced). */
int nefct_backproject_cyli(const float* projection_value, const int grid[3], const float center[3],
                           const float size[3], const float* angles, int nv, float SID, float SAD,
                           int na, float da, float ai, int nb, float db, float bi, float* image,
                           unsigned flags, void* workspace, size_t workspace_bytes, void* stream);

/* Helical forward projection and adjoint: as nefct_project_flat/_cyli and
 * nefct_backproject_flat/_cyli (cyl = 0 flat, 1 cylindrical), with view iv's
 * source and detector lifted by offsets[iv] mm along z (device array of nv
 * floats, required).  Replaces the per-view 'offsets' the reference's Python
 * callers pass to their projector (nefct/functions/project.py:74-76,
 * back_project.py:157, project_t.py:157), with the lift of
 * dist_3d_back_fdk.cu.cc:46,65.  The pair stays an exact transpose. */
int nefct_project_helical(const float* image, const int grid[3], const float center[3], const float size[3],
                          const float* angles, const float* offsets, int nv, float SID, float SAD, int na,
                          float da, float ai, int nb, float db, float bi, int cyl, float* projection_value,
                          unsigned flags, void* workspace, size_t workspace_bytes, void* stream);
int nefct_backproject_helical(const float* projection_value, const int grid[3], const float center[3],
                              const float size[3], const float* angles, const float* offsets, int nv, float SID,
                              float SAD, int na, float da, float ai, int nb, float db, float bi, int cyl,
                              float* image, unsigned flags, void* workspace, size_t workspace_bytes, void* stream);

/* Per-ray count of voxels with nonzero length (N_int) as walked by the exact
 * kernels; counts {nv, nb, na} int32.  No reference counterpart: it is the
 * numerator of the bandwidth figures (voxel reads per ray). */
int nefct_siddon_count(const int grid[3], const float center[3], const float size[3],
                       const float* angles, int nv, float SID, float SAD, int na, float da, float ai,
                       int nb, float db, float bi, int cyl, int32_t* counts, unsigned flags,
                       void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched Siddon pair: many problems of uneven size in one launch per     */
/* direction.            No reference counterpart: it replaces the        */
/* reference's per-view Python loops (nefct/functions/project_t.py:153-159, */
/* back_project_t.py:150-156) and per-problem launches.                     */
/* ---------------------------------------------------------------------- */
typedef struct {
    /* set by the caller (device pointers, siddon-op geometry) */
    const float* volume;  /* forward input / adjoint output, x fastest */
    float* sinogram;      /* forward output / adjoint input, [nv][nb][na] */
    const float* angles;  /* [nv] */
    int grid[3];
    float center[3];
    float size[3];
    int nv;
    int cyl;
    float SID, SAD;
    int na, nb;
    float da, ai, db, bi;
    /* filled by nefct_batch_plan */
    long long tile_begin;   /* first forward task (16x16 rays of one view; views padded to groups of 16) */
    long long block_begin;  /* first adjoint task (one 32x32x16 voxel block) */
    long long ws_offset;    /* byte offset of the problem's workspace region */
    int tiles_a, tiles_b;
    int blocks[3];
    int pad_;
} nefct_problem;

/* Validates the problems, fills their plan fields and returns the workspace
 * size and task totals.  Host-only; no device work. */
int nefct_batch_plan(nefct_problem* problems, int nprob, unsigned flags, size_t* workspace_bytes,
                     long long* total_tiles, long long* total_blocks);

/* Forward projection of every problem: one trig/launch-record pass, the
 * padded volume copies, then one launch over the 16x16-ray tasks of all
 * problems, dealt to the XCDs in contiguous ranges.  'host_problems' is the
 * planned table, 'device_problems' a device copy of it. */
int nefct_batch_project(const nefct_problem* host_problems, const nefct_problem* device_problems, int nprob,
                        long long total_tiles, unsigned flags, void* workspace, size_t workspace_bytes, void* stream);

/* Exact (or reference-mode) adjoint of every problem: one max-|projection|
 * pass, then one launch with a workgroup per 32x32x16 voxel block.  Every
 * output volume is fully written.  NEFCT_F_LANE_STRIDE2 spaces a wavefront's
 * lanes two detector rows apart, as in the plain adjoint (identical result);
 * the other launch-shape flags do not apply to the batched kernels. */
int nefct_batch_backproject(const nefct_problem* host_problems, const nefct_problem* device_problems, int nprob,
                            long long total_blocks, unsigned flags, void* workspace, size_t workspace_bytes,
                            void* stream);

/* ---------------------------------------------------------------------- */
/* Trilinear warp                                                          */
/* ---------------------------------------------------------------------- */

/* Pull-trilinear warp: out[i] = image(i + (mx,my,mz)[i]), displacement in
 * voxel units, all volumes x fastest with the same grid.  boundary 0
 * reproduces the reference: a voxel is written only when every base index
 * floor(p) is in [1, n-2] (cpp/deform/deform.cu.cc:31-33), else 0 (the output
 * is zero-filled as in nefct/cc/deform/deform.cc:94); boundary 1 = zero
 * padding (each in-range corner contributes).  Replaces DeformOp::Compute
 * (cpp/deform/deform.cc:64-90) + deform (cpp/deform/deform.cu.cc:112-126). */
int nefct_deform(const float* image, const float* mx, const float* my, const float* mz, const int grid[3],
                 float* deformed_image, int boundary, void* stream);

/* Transpose of nefct_deform in the image (scatter of grad_out to the 8
 * corners).  grad_image is fully written.  No reference counterpart (the
 * reference registers no gradient); it is the backward of the warp. */
int nefct_deform_adjoint(const float* grad_out, const float* mx, const float* my, const float* mz,
                         const int grid[3], float* grad_image, int boundary, void* stream);

/* Gradient of <grad_out, deform(image, m)> with respect to mx, my, mz
 * (voxel units).  No reference counterpart; backward of the warp. */
int nefct_deform_dvf_grad(const float* image, const float* grad_out, const float* mx, const float* my,
                          const float* mz, const int grid[3], float* grad_mx, float* grad_my, float* grad_mz,
                          int boundary, void* stream);

/* ---------------------------------------------------------------------- */
/* Texture-semantics warps (nefct/cc/deform/deform_tex.{cc,cu.cc})         */
/* ---------------------------------------------------------------------- */
/* Trilinear sampling with the reference's texture rule (deform_tex.cu.cc:60-69:
 * unnormalised coordinates, linear filter, clamp addressing): along each axis
 * coordinate u covers texel i on [i, i+1); value (1-a) T[clamp(i0)] +
 * a T[clamp(i0+1)], i0 = floor(u - 1/2), a = u - 1/2 - i0, clamp to [0, n-1].
 * Weights are exact FP32 (texture hardware rounds a to 1/256).  Volumes x
 * fastest, displacements in voxel units; no allocation, no synchronisation. */

/* out[i] = T(i + m(i) + 1/2).  Replaces deform_tex (deform_tex.cu.cc:29-76,
 * DeformKernel :15-26) behind op Deform (deform_tex.cc:12-22). */
int nefct_deform_tex(const float* image, const float* mx, const float* my, const float* mz, const int grid[3],
                     float* deformed_image, void* stream);

/* Fixed-point inverse of a displacement field: per voxel d = 0, then
 * 'iterations' times x = -mx(i + d), y = -my(i + d), z = -mz(i + d), each
 * sample using the components already updated.  Replaces invert
 * (deform_tex.cu.cc:102-168, DeformInvertKernel :78-100; 30 iterations there). */
int nefct_invert_dvf(const float* mx, const float* my, const float* mz, const int grid[3], int iterations,
                     float* inv_mx, float* inv_my, float* inv_mz, void* stream);

/* Warp by the inverted field (30 iterations).  workspace >= 12 * voxels bytes
 * (the inverted field).  Replaces deform_invert_tex (deform_tex.cu.cc:170-226)
 * behind op DeformInvert (deform_tex.cc:24-34). */
int nefct_deform_invert_tex(const float* image, const float* mx, const float* my, const float* mz, const int grid[3],
                            float* deformed_image, void* workspace, size_t workspace_bytes, void* stream);

/* out[i] = T(i - (mx, my, mz) + 1/2).  Replaces shift_tex (deform_tex.cu.cc:
 * 242-289, ShiftKernel :228-239) behind op Shift (deform_tex.cc:36-46). */
int nefct_shift_tex(const float* image, const int grid[3], float mx, float my, float mz, float* deformed_image,
                    void* stream);

/* Shift along z only.  Replaces shift_z_tex (deform_tex.cu.cc:306-353,
 * ShiftZKernel :291-302) behind op ShiftZ (deform_tex.cc:48-56). */
int nefct_shift_z_tex(const float* image, const int grid[3], float mz, float* deformed_image, void* stream);

/* In-plane rotation about the integer centre (nx/2, ny/2): sample at
 * ((ix+1/2-nx/2) cos - (iy+1/2-ny/2) sin + nx/2, (ix+1/2-nx/2) sin +
 * (iy+1/2-ny/2) cos + ny/2, iz + 1/2).  Replaces rotate_tex
 * (deform_tex.cu.cc:373-420, RotateKernel :355-370) behind op Rotate
 * (deform_tex.cc:58-66). */
int nefct_rotate_tex(const float* image, const int grid[3], float cos_phi, float sin_phi, float* deformed_image,
                     void* stream);

/* ---------------------------------------------------------------------- */
/* FDK (flat detector)                                                     */
/* ---------------------------------------------------------------------- */

/* out = projection_value * cosine weight * [Parker weight] * scale.
 * Cosine weight SAD/sqrt(SAD^2 + a^2 + b^2) (mat_codes/filtering.m:5-14),
 * Parker short-scan weights with views at ((180/dang - nv)/2 + i + 1) dang
 * degrees (mat_codes/ParkerWeight.m:5-27); direction +1/-1 is the rotation
 * sense.  No reference CUDA counterpart (the reference does this in MATLAB). */
int nefct_fdk_preweight(const float* projection_value, float* out, int nv, float SID, float SAD, int na, float da,
                        float ai, int nb, float db, float bi, int parker, float dang_deg, int direction, float scale,
                        void* stream);

/* Row filtering: out[r][u] = sum_v in[r][v] taps[u - v + na - 1] for rows r
 * of length na (<= 2048), taps of length 2 na - 1 (the spatial form of the
 * windowed ramp of mat_codes/filtering.m:20-41,54-88; in != out). */
int nefct_fdk_filter(const float* in, float* out, int rows, int na, const float* taps, void* stream);

/* Workspace bytes of nefct_fdk_backproject (per-view table). */
size_t nefct_fdk_workspace_size(int nv);

/* Voxel-driven FDK backprojection: image[v] = sum over views of
 * (SID/U)^2 * bilinear(filtered view at the projection of voxel centre v),
 * U = SID + x cos t + y sin t; offsets[iv] (may be NULL) shifts source and
 * detector along z (helical), as dist_3d_back_fdk.cu.cc:44-65.  Image fully
 * written.  Replaces back_project_flat_fdk_gpu (dist_3d_back_fdk.cu.cc:257-276)
 * with the voxel-driven sampling of mat_codes/backprojection.m:11-25. */
int nefct_fdk_backproject(const float* filtered, const int grid[3], const float center[3], const float size[3],
                          const float* angles, const float* offsets, int nv, float SID, float SAD, int na, float da,
                          float ai, int nb, float db, float bi, float* image, void* workspace,
                          size_t workspace_bytes, void* stream);

/* nefct_fdk_backproject with a chosen work mapping (identical images):
 * NEFCT_FDK_BP_AUTO the 16x16x16 staged mapping where eligible, else gathers
 * (what nefct_fdk_backproject does); NEFCT_FDK_BP_GATHER every voxel's taps gathered
 * from global memory; NEFCT_FDK_BP_STAGED each 16x16x16-voxel block's detector
 * footprint staged in LDS once per view; NEFCT_FDK_BP_WIDE the same for
 * 64x64x16-voxel blocks (a fifth of the HBM traffic, slower).  The staged
 * mappings return NEFCT_ERR_UNSUPPORTED unless na, nb >= 2 and every voxel
 * centre lies within 0.999 SID of the rotation axis in-plane.  No reference
 * counterpart (a tuning choice). */
#define NEFCT_FDK_BP_AUTO 0
#define NEFCT_FDK_BP_GATHER 1
#define NEFCT_FDK_BP_STAGED 2
#define NEFCT_FDK_BP_WIDE 3
int nefct_fdk_backproject_mapped(const float* filtered, const int grid[3], const float center[3],
                                 const float size[3], const float* angles, const float* offsets, int nv, float SID,
                                 float SAD, int na, float da, float ai, int nb, float db, float bi, float* image,
                                 int mapping, void* workspace, size_t workspace_bytes, void* stream);

/* Voxel-driven backprojection without weights: the reference's second
 * BackProjectFlat / BackProjectCyli op pair (nefct/cc/pixel_3d/pixel.cc:11-47,
 * launchers pixel.cu.cc:126-204, kernels :12-124).  image[v] = sum over views
 * of the bilinear sample of projection_value at the central projection of
 * voxel centre v; views whose 2 x 2 neighbourhood leaves the detector are
 * skipped.  The reference's conventions define the op: voxel centre
 * (i + 0.5) size/n + center (no -size/2), source at distance SID from the
 * axis, detector at SAD from the source, index y/d + n/2 (integer n/2)
 * - offset - 0.5.  Host grid/center/size; device projection [nv][nb][na],
 * angles [nv] (radians), image [nz][ny][nx] (fully written).  Limits:
 * 4 na < 2^24, nb < 2^24, 4 na nb < 2^32 (else NEFCT_ERR_SHAPE).  Volumes
 * with every voxel centre in front of the source run on the LDS-staged blocks
 * of nefct_fdk_backproject_mapped (rows and flat columns in FP64, cylindrical
 * columns by FP32 atan2); the rest use the op's FP32 arithmetic. */
int nefct_pixel_backproject_flat(const float* projection_value, const int grid[3], const float center[3],
                                 const float size[3], const float* angles, int nv, float SID, float SAD, int na,
                                 float da, float ai, int nb, float db, float bi, float* image, void* stream);
int nefct_pixel_backproject_cyli(const float* projection_value, const int grid[3], const float center[3],
                                 const float size[3], const float* angles, int nv, float SID, float SAD, int na,
                                 float da, float ai, int nb, float db, float bi, float* image, void* stream);

/* Pixel-driven FDK backprojection with slab footprints: the arithmetic of
 * nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255; replaces its launchers
 * back_project_flat_fdk_gpu (:257-276) and back_project_cyli_fdk_gpu
 * (:278-298), same argument order.  That file's units and naming: lengths in
 * voxel units, volume centred on the axis, SAD = source-to-axis distance,
 * SID = source-to-detector distance, da in voxel units (flat) or radians
 * (cylindrical).  shape = (nx, ny, nz) on the host (the reference copies it
 * from the device); offsets[nv] (may be NULL = 0) lift source and detector
 * along z.  Adds into image [nz][ny][nx], as the reference does (the caller
 * zeroes it first); additions are atomic, so the summation order varies.
 * A workgroup (64 columns x 4 rows of one view) first sums its additions over
 * a few slabs in 64-bit fixed point whose step is 2^-30 of its largest
 * possible contribution (max |pv| |r|/|r_dom| SAD^2 / U_min^2). */
int nefct_fdk_footprint_backproject_flat(const float* pv_values, const int shape[3], const float* offsets,
                                         const float* angles, int nv, float SID, float SAD, float da, float ai,
                                         int na, float db, float bi, int nb, float* image, void* stream);
int nefct_fdk_footprint_backproject_cyli(const float* pv_values, const int shape[3], const float* offsets,
                                         const float* angles, int nv, float SID, float SAD, float da, float ai,
                                         int na, float db, float bi, int nb, float* image, void* stream);

/* Smoothed anisotropic total variation TV(x) = sum_k sum_v sqrt(d_k(v)^2 +
 * eps^2) over the forward differences with a zero first slice of
 * nefct/functions/total_variation.py:19-35 (whose transpose is
 * InvertTotalVari, :125-149).  grad (may be NULL) = scale * dTV/dx, fully
 * written; *value (may be NULL; device FP64) += scale * TV(x).  image [nz][ny][nx].
 * No reference counterpart for the smoothing (the reference's TV is used
 * inside ADMM, conj_grad.py:20-24). */
int nefct_tv_smooth(const float* image, const int grid[3], float eps, float scale, float* grad, double* value,
                    void* stream);

/* *value (device FP64) += scale * sum x[i]^2, accumulated in FP64; the
 * caller clears *value first.  For loss values inside captured graphs. */
int nefct_sq_norm(const float* x, size_t n, double scale, double* value, void* stream);

/* One fused gradient-descent update of 0.5 ||A x - b||^2 + lam TV(x):
 * out = image - step * (data_grad + lam * dTV/dx(image)), data_grad =
 * A^T (A x - b) from the adjoint.  out != image.  No reference counterpart
 * (the reference iterates in numpy, nefct/functions/sart.py:29-39,
 * conj_grad.py:50-62). */
int nefct_tv_descent_step(const float* image, const float* data_grad, const int grid[3], float eps, float lam,
                          float step, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NEFCT_HIP_H */
