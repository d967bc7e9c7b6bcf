// Pixel-driven FDK backprojection with slab footprints for gfx950: the
// arithmetic of nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255 (the
// reference's only CUDA FDK backprojector; it has no op registration, its
// launchers are back_project_{flat,cyli}_fdk_gpu, :257-298).
//
// Units and naming are that file's: lengths in voxel units with the volume
// centred on the rotation axis, SAD = source-to-axis distance, SID =
// source-to-detector distance, helical offsets[iv] lift source and detector.
// For each pixel and each slab of the pixel's dominant in-plane axis, the
// footprint on the slab mid-plane is bounded in the transverse in-plane axis
// by the hull of the pixel's two edge rays (all four slope pairings, :104-125)
// and in z by the centre ray +- db/2 (:127-138); every overlapped voxel gets
//   pv * wt * wz * |r| / |r_dom| * SAD^2 / (SAD + x cos + y sin)^2   (:142-157)
// with wt, wz the overlap fractions of the footprint.  The integer n/2 of the
// reference's index shifts is kept.
//
// Mapping: one lane per pixel, a wavefront = 64 neighbouring columns of one
// detector row, so at each slab the lanes' footprints are neighbours and
// their atomic additions meet in the same L2 lines.  The per-pixel edge rays,
// slopes, |r| and z bounds are computed once per pixel (the reference
// recomputes the edge rays, with cosf/sinf for the cylinder, in every slab).
// Additions go through buffer atomics whose range check covers the volume.
#include <math.h>

#include <cmath>
#include <initializer_list>

#include "common.h"

namespace {

// floor(v) clamped to [lo, hi] without relying on float->int saturation
__device__ __forceinline__ int floor_clamp(float v, int lo, int hi)
{
    const float f = floorf(v);
    if (!(f >= (float)lo)) return lo;
    if (f > (float)hi) return hi;
    return (int)f;
}

template <bool CYL>
__device__ __forceinline__ void edge_point(float SID, float SAD, float da, float ai, int na, float e, float c,
                                           float s, float x1, float y1, float& xo, float& yo)
{
    float x20, y20;
    const float ang_or_a = e * da + ai - na * da / 2;
    if (CYL) {
        x20 = SID * cosf(ang_or_a) - SAD;
        y20 = SID * sinf(ang_or_a);
    } else {
        x20 = SID - SAD;
        y20 = ang_or_a;
    }
    xo = x20 * c - y20 * s - x1;
    yo = x20 * s + y20 * c - y1;
}

template <bool CYL>
__global__ __launch_bounds__(256) void k_fdk_fp_bp(const float* __restrict__ pv, const float* __restrict__ angles,
                                                   const float* __restrict__ offsets, int nv, float SID, float SAD,
                                                   int nx, int ny, int nz, float da, float ai, int na, float db,
                                                   float bi, int nb, float* __restrict__ image)
{
    const int ia = blockIdx.x * 64 + (threadIdx.x & 63);
    const int ib = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iv = blockIdx.z;
    if (ia >= na || ib >= nb) return;
    const float val = pv[(size_t)ia + (size_t)na * ((size_t)ib + (size_t)nb * iv)];
    if (val == 0.0f) return;   // adds nothing
    const double ang = (double)angles[iv];
    const float c = (float)cos(ang), s = (float)sin(ang);
    const float off = offsets ? offsets[iv] : 0.0f;
    const float x1 = -SAD * c, y1 = -SAD * s, z1 = off;
    float x21, y21;
    edge_point<CYL>(SID, SAD, da, ai, na, (float)ia + 0.5f, c, s, x1, y1, x21, y21);
    const float z21 = ((ib + 0.5f) * db + bi - nb * db / 2 + off) - z1;
    float xl, yl, xr, yr;
    edge_point<CYL>(SID, SAD, da, ai, na, (float)ia, c, s, x1, y1, xl, yl);
    edge_point<CYL>(SID, SAD, da, ai, na, (float)(ia + 1), c, s, x1, y1, xr, yr);
    const bool xdom = fabsf(x21) > fabsf(y21);
    const float len = sqrtf(x21 * x21 + y21 * y21 + z21 * z21);
    // walk axis m (x or y), transverse in-plane axis t
    const int nm = xdom ? nx : ny, nt = xdom ? ny : nx;
    const float rdom = xdom ? x21 : y21;
    float q[4];
    if (xdom) {
        q[0] = yl / xl; q[1] = yl / xr; q[2] = yr / xl; q[3] = yr / xr;
    } else {
        q[0] = xl / yl; q[1] = xl / yr; q[2] = xr / yl; q[3] = xr / yr;
    }
    const float src_m = xdom ? x1 : y1;
    const float base_t = (xdom ? y1 : x1) + (float)((xdom ? ny : nx) / 2);
    const float kzl = (z21 - db / 2) / rdom, kzr = (z21 + db / 2) / rdom;
    const float base_z = z1 + (float)(nz / 2);
    const float half_m = (float)nm / 2;
    const int hx = nx / 2, hy = ny / 2;
    const float pre = len / fabsf(rdom) * SAD * SAD;
    const __amdgpu_buffer_rsrc_t img = make_rsrc(image, (uint32_t)((size_t)nx * ny * nz * 4));
    for (int im = 0; im < nm; ++im) {
        const float mc = (float)im + 0.5f - half_m - src_m;
        float t1 = q[0] * mc + base_t, t2 = t1;
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            const float v = q[k] * mc + base_t;
            t1 = fminf(t1, v);
            t2 = fmaxf(t2, v);
        }
        const float za = kzl * mc + base_z, zb = kzr * mc + base_z;
        const float z1f = fminf(za, zb), z2f = fmaxf(za, zb);
        const int it0 = floor_clamp(t1, 0, nt), it1 = floor_clamp(t2, -1, nt - 1);
        const int iz0 = floor_clamp(z1f, 0, nz), iz1 = floor_clamp(z2f, -1, nz - 1);
        if (it0 > it1 || iz0 > iz1) continue;
        const float inv_t = 1.0f / (t2 - t1), inv_z = 1.0f / (z2f - z1f);
        for (int it = it0; it <= it1; ++it) {
            const float wt = (fminf(it + 1.0f, t2) - fmaxf((float)it, t1)) * inv_t;
            const int ix = xdom ? im : it, iy = xdom ? it : im;
            const float xc = ix + 0.5f - hx, yc = iy + 0.5f - hy;
            const float u = SAD + xc * c + yc * s;
            const float w_t = val * wt * pre / (u * u);
            for (int iz = iz0; iz <= iz1; ++iz) {
                const float wz = (fminf(iz + 1.0f, z2f) - fmaxf((float)iz, z1f)) * inv_z;
                const int o = ix + nx * (iy + ny * iz);
                buf_atomic_add(img, o * 4, w_t * wz);
            }
        }
    }
}

bool finite_all(std::initializer_list<float> v)
{
    for (float x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int run(bool cyl, const float* pv, const int shape[3], const float* offsets, const float* angles, int nv, float SID,
        float SAD, float da, float ai, int na, float db, float bi, int nb, float* image, void* stream)
{
    if (!pv || !shape || !angles || !image) return NEFCT_ERR_NULL;
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    if (nx <= 0 || ny <= 0 || nz <= 0 || nv <= 0 || na <= 0 || nb <= 0) return NEFCT_ERR_SHAPE;
    if (nv > 65535 || (nb + 3) / 4 > 65535 || (size_t)nx * ny * nz * 4 >= 0x80000000ull) return NEFCT_ERR_SHAPE;
    if (!finite_all({SID, SAD, da, ai, db, bi}) || !(SAD > 0 && SID > 0 && da > 0 && db > 0))
        return NEFCT_ERR_GEOMETRY;
    const dim3 g((na + 63) / 64, (nb + 3) / 4, nv);
    hipStream_t s = (hipStream_t)stream;
    if (cyl)
        hipLaunchKernelGGL(k_fdk_fp_bp<true>, g, dim3(256), 0, s, pv, angles, offsets, nv, SID, SAD, nx, ny, nz, da,
                           ai, na, db, bi, nb, image);
    else
        hipLaunchKernelGGL(k_fdk_fp_bp<false>, g, dim3(256), 0, s, pv, angles, offsets, nv, SID, SAD, nx, ny, nz, da,
                           ai, na, db, bi, nb, image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // namespace

extern "C" {

int nefct_fdk_footprint_backproject_flat(const float* pv_values, const int shape[3], const float* offsets,
                                         const float* angles, int nv, float SID, float SAD, float da, float ai,
                                         int na, float db, float bi, int nb, float* image, void* stream)
{
    return run(false, pv_values, shape, offsets, angles, nv, SID, SAD, da, ai, na, db, bi, nb, image, stream);
}

int nefct_fdk_footprint_backproject_cyli(const float* pv_values, const int shape[3], const float* offsets,
                                         const float* angles, int nv, float SID, float SAD, float da, float ai,
                                         int na, float db, float bi, int nb, float* image, void* stream)
{
    return run(true, pv_values, shape, offsets, angles, nv, SID, SAD, da, ai, na, db, bi, nb, image, stream);
}

}  // extern "C"
