// Voxel-driven backprojection without weights: the reference's second pair
// of BackProjectFlat / BackProjectCyli ops (nefct/cc/pixel_3d/pixel.cc:11-47,
// kernels nefct/cc/pixel_3d/pixel.cu.cc:12-124), rebuilt for gfx950.
//
// Per voxel and view the voxel centre is rotated into the view frame, its
// detector coordinates are found by central projection from the source, and
// the projection is sampled bilinearly; views whose 2 x 2 neighbourhood is
// not fully on the detector are skipped.  The reference's conventions are
// kept because they define the op:
//   voxel centre  x = (i + 0.5) size / n + center        (no -size/2, pixel.cu.cc:116-118)
//   flat          a = (SAD / (x' + SID)) y' / da + na/2 - ai - 0.5   (integer na/2, :31-34)
//   cylindrical   a = atan2(y', x' + SID) / da + na/2 - ai - 0.5     (:70-71)
//   both          b = (SAD / (x' + SID)) z / db + nb/2 - bi - 0.5    (:72-73)
// with (x', y') the voxel centre rotated by -angle and the source at
// (-SID, 0, 0) of that frame.  The product is accumulated in FP32 over the
// views in order, as the reference does (pixel.cu.cc:52).
//
// Calls go to the LDS-staged mapping of fdk.hip (16 x 16 x 16 voxel blocks,
// each view's block footprint copied to LDS once; detector rows in FP64,
// flat columns in FP64 and cylindrical ones by FP32 atan2; 1.7x faster at
// 512 x 512 x 256 voxels and 408 views of 1024 x 768) wherever every voxel
// centre lies in front of the source; the kernel below takes the rest, with
// the op's FP32 arithmetic.
//
// Layout and launch: one lane per (x, y) column of voxels, a wavefront spans
// 64 consecutive x, each lane keeps PZ z-values in registers, so the
// per-view projection (x', y', magnification, a) is computed once per column
// and reused for PZ voxels.  cos/sin of a chunk of views are staged in LDS
// by the workgroup (computed in FP64 and rounded to FP32).  Per voxel the row
// coordinate is one FMA, b = z * (t / db) + (nb/2 - bi - 0.5), instead of
// the reference's product, division and two subtractions (the same value to
// an ulp or two of b), and the four taps are two 8-byte loads (the skip rule
// keeps the 2 x 2 neighbourhood on the detector, so ia <= na - 2 and
// ib <= nb - 2), issued for all PZ voxels before their arithmetic; a voxel
// whose neighbourhood leaves the detector is clamped for the address and its
// sample dropped by a select.  The bilinear sample is formed as two lerps.
// The taps of neighbouring lanes are neighbouring detector columns of the
// same view, so the gathers coalesce.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "fdk_staged.h"
#include "nefct_hip.h"

namespace {

constexpr int PZ = 16;     // z-values per lane
constexpr int PV = 256;    // views staged in LDS per chunk

template <bool CYL>
__global__ __launch_bounds__(256) void k_pixel_bp(const float* __restrict__ pv, const float* __restrict__ angles,
                                                  int nv, int nx, int ny, int nz, float cx, float cy, float cz,
                                                  float sx, float sy, float sz, float SID, float SAD, int na,
                                                  float da, float ai, int nb, float db, float bi,
                                                  float* __restrict__ image)
{
    __shared__ float s_c[PV], s_s[PV];
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz0 = blockIdx.z * PZ;
    const bool live = ix < nx && iy < ny;
    const int kn = min(PZ, nz - iz0);
    // voxel centres, formed in double and rounded as the reference does
    const float xc = (float)((ix + 0.5) * sx / nx + cx);
    const float yc = (float)((iy + 0.5) * sy / ny + cy);
    float zc[PZ];
#pragma unroll
    for (int k = 0; k < PZ; ++k) zc[k] = (float)((iz0 + k + 0.5) * sz / nz + cz);
    float acc[PZ];
#pragma unroll
    for (int k = 0; k < PZ; ++k) acc[k] = 0.0f;
    const float ha = (float)(na / 2), hb = (float)(nb / 2);
    const float cb = hb - bi - 0.5f, nb1 = (float)(nb - 1);
    const unsigned rowstep = 4u * (unsigned)na;
    const unsigned row1 = nb >= 2 ? rowstep : 0u;   // one-row detectors: every voxel is skipped
    for (int v0 = 0; v0 < nv; v0 += PV) {
        const int nvc = min(PV, nv - v0);
        __syncthreads();
        if ((int)threadIdx.x < nvc) {
            const double a = (double)angles[v0 + threadIdx.x];
            s_c[threadIdx.x] = (float)cos(a);
            s_s[threadIdx.x] = (float)sin(a);
        }
        __syncthreads();
        if (!live) continue;
        for (int j = 0; j < nvc; ++j) {
            const float c = s_c[j], s = s_s[j];
            const float xc0 = xc * c + yc * s;
            const float yc0 = -xc * s + yc * c;
            const float t = SAD / (xc0 + SID);
            float a;
            if (CYL) {
                a = atan2f(yc0, xc0 + SID) / da + ha - ai - 0.5f;
            } else {
                const float yd = t * yc0;
                a = yd / da + ha - ai - 0.5f;
            }
            // ia = floor(a) in [0, na - 2] (pixel.cu.cc:43-46); NaN fails too
            if (!(a >= 0.0f && a < (float)(na - 1))) continue;
            const float fa = floorf(a);
            const int ia = (int)fa;
            const float wa2 = a - fa;
            const float tdb = t / db;
            const char* pview = reinterpret_cast<const char*>(pv + (size_t)(v0 + j) * na * nb) + 4u * (unsigned)ia;
            float2 p0[PZ], p1[PZ];
            float wb2[PZ];
            bool valid[PZ];
            // every voxel of a full chunk on the detector (b is monotone in z,
            // so its ends bound it): no clamp, no validity test, no select
            const float bf = fmaf(zc[0], tdb, cb), bl = fmaf(zc[PZ - 1], tdb, cb);
            if (kn == PZ && fminf(bf, bl) >= 0.0f && fmaxf(bf, bl) < nb1) {
#pragma unroll
                for (int k = 0; k < PZ; ++k) {
                    const float b = fmaf(zc[k], tdb, cb);
                    const int ib = (int)b;                  // 0 <= b < nb - 1
                    wb2[k] = b - (float)ib;
                    const unsigned o0 = __umul24((unsigned)ib, rowstep);
                    __builtin_memcpy(&p0[k], pview + o0, 8);
                    __builtin_memcpy(&p1[k], pview + o0 + row1, 8);
                }
#pragma unroll
                for (int k = 0; k < PZ; ++k) {
                    const float top = fmaf(wa2, p0[k].y - p0[k].x, p0[k].x);
                    const float bot = fmaf(wa2, p1[k].y - p1[k].x, p1[k].x);
                    acc[k] = acc[k] + fmaf(wb2[k], bot - top, top);
                }
                continue;
            }
#pragma unroll
            for (int k = 0; k < PZ; ++k) {
                const float b = fmaf(zc[k], tdb, cb);
                valid[k] = k < kn && b >= 0.0f && b < nb1;   // (pixel.cu.cc:47-50)
                const float bcl = __builtin_amdgcn_fmed3f(b, 0.0f, nb1);
                const int ib = min((int)bcl, max(nb - 2, 0));   // bcl >= 0: truncation is floor
                wb2[k] = bcl - (float)ib;
                const unsigned o0 = __umul24((unsigned)ib, rowstep);
                __builtin_memcpy(&p0[k], pview + o0, 8);
                __builtin_memcpy(&p1[k], pview + o0 + row1, 8);
            }
#pragma unroll
            for (int k = 0; k < PZ; ++k) {
                const float top = fmaf(wa2, p0[k].y - p0[k].x, p0[k].x);
                const float bot = fmaf(wa2, p1[k].y - p1[k].x, p1[k].x);
                const float sum = acc[k] + fmaf(wb2[k], bot - top, top);
                acc[k] = valid[k] ? sum : acc[k];
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int k = 0; k < PZ; ++k) {
        const int iz = iz0 + k;
        if (iz < nz) image[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)] = acc[k];
    }
}

int run(bool cyl, const float* pv, const int grid[3], const float center[3], const float size[3],
        const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb, float db, float bi,
        float* image, void* stream)
{
    if (!pv || !grid || !center || !size || !angles || !image) return NEFCT_ERR_NULL;
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nv <= 0 || na <= 0 || nb <= 0) return NEFCT_ERR_SHAPE;
    if ((ny + 3) / 4 > 65535 || (nz + PZ - 1) / PZ > 65535) return NEFCT_ERR_SHAPE;
    // in-view byte offsets are 24-bit products (row * 4 na) and 32-bit sums
    if (4ull * na >= (1ull << 24) || nb >= (1 << 24) || 4ull * na * nb >= (1ull << 32)) return NEFCT_ERR_SHAPE;
    for (float v : {center[0], center[1], center[2], size[0], size[1], size[2], SID, SAD, da, ai, db, bi})
        if (!std::isfinite(v)) return NEFCT_ERR_GEOMETRY;
    if (!(size[0] > 0 && size[1] > 0 && size[2] > 0 && da > 0 && db > 0 && SID > 0 && SAD > 0))
        return NEFCT_ERR_GEOMETRY;
    hipStream_t s = (hipStream_t)stream;
    // the LDS-staged mapping where the geometry allows it
    if (nefct_internal::pixel_bp_staged(cyl, pv, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi,
                                        image, s)) {
        NEFCT_CHECK_LAUNCH();
        return NEFCT_OK;
    }
    const dim3 g((nx + 63) / 64, (ny + 3) / 4, (nz + PZ - 1) / PZ);
    if (cyl)
        hipLaunchKernelGGL(k_pixel_bp<true>, g, dim3(256), 0, s, pv, angles, nv, nx, ny, nz, center[0], center[1],
                           center[2], size[0], size[1], size[2], SID, SAD, na, da, ai, nb, db, bi, image);
    else
        hipLaunchKernelGGL(k_pixel_bp<false>, g, dim3(256), 0, s, pv, angles, nv, nx, ny, nz, center[0], center[1],
                           center[2], size[0], size[1], size[2], SID, SAD, na, da, ai, nb, db, bi, image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // namespace

extern "C" {

int nefct_pixel_backproject_flat(const float* projection_value, const int grid[3], const float center[3],
                                 const float size[3], const float* angles, int nv, float SID, float SAD, int na,
                                 float da, float ai, int nb, float db, float bi, float* image, void* stream)
{
    return run(false, projection_value, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, image,
               stream);
}

int nefct_pixel_backproject_cyli(const float* projection_value, const int grid[3], const float center[3],
                                 const float size[3], const float* angles, int nv, float SID, float SAD, int na,
                                 float da, float ai, int nb, float db, float bi, float* image, void* stream)
{
    return run(true, projection_value, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, image,
               stream);
}

}  // extern "C"
