// Texture-semantics warps for gfx950: dense-field warp, fixed-point inverse of
// a displacement field, constant shift and in-plane rotation.
//
// Replaces nefct/cc/deform/deform_tex.cu.cc (DeformKernel :15-26,
// DeformInvertKernel :78-100, ShiftKernel :228-239, ShiftZKernel :291-302,
// RotateKernel :355-370 and their host launchers) behind the ops of
// deform_tex.cc:12-63.  The reference samples a 3-D texture (unnormalised
// coordinates, linear filtering, clamp addressing) built per call from a copy
// of the volume; here the same trilinear rule is evaluated from the volume in
// HBM: along each axis coordinate u covers texel i on [i, i + 1), so the value
// is (1 - a) T[clamp(i0)] + a T[clamp(i0 + 1)] with i0 = floor(u - 1/2),
// a = u - 1/2 - i0, and indices clamped to [0, n - 1].  The weights are exact
// FP32 (texture hardware quantises a to 1/256).
//
// One lane per output voxel, x fastest: neighbouring lanes sample neighbouring
// points, so the 8 gathers of a wavefront fall in a few cache lines (L2
// bound; HBM traffic is one read of the inputs and one write).
#include <cmath>

#include "common.h"

namespace {

struct Axis {
    int i0, i1;
    float a;
};

// texel pair and weight along one axis; NaN coordinates give NaN weights, huge
// ones clamp to the edge texel
__device__ __forceinline__ Axis axis_of(float u, int n)
{
    Axis r;
    const float xb = u - 0.5f;
    const float fl = floorf(xb);
    r.a = xb - fl;
    const int i = (int)fminf(fmaxf(fl, -1.0f), (float)n - 1.0f);
    r.i0 = max(i, 0);
    r.i1 = min(i + 1, n - 1);
    if (!(fabsf(xb) < 1e30f)) r.a = (xb == xb) ? 0.0f : r.a;   // +-inf: edge texel exactly
    return r;
}

__device__ __forceinline__ float tex3(const float* __restrict__ v, int nx, int ny, int nz, float u, float w, float t)
{
    const Axis X = axis_of(u, nx), Y = axis_of(w, ny), Z = axis_of(t, nz);
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const float* p00 = v + sy * Y.i0 + sz * Z.i0;
    const float* p10 = v + sy * Y.i1 + sz * Z.i0;
    const float* p01 = v + sy * Y.i0 + sz * Z.i1;
    const float* p11 = v + sy * Y.i1 + sz * Z.i1;
    const float bx = 1.0f - X.a, by = 1.0f - Y.a, bz = 1.0f - Z.a;
    const float c00 = bx * p00[X.i0] + X.a * p00[X.i1];
    const float c10 = bx * p10[X.i0] + X.a * p10[X.i1];
    const float c01 = bx * p01[X.i0] + X.a * p01[X.i1];
    const float c11 = bx * p11[X.i0] + X.a * p11[X.i1];
    return bz * (by * c00 + Y.a * c10) + Z.a * (by * c01 + Y.a * c11);
}

// voxel of linear index id (the volume has fewer than 2^29 voxels)
__device__ __forceinline__ bool voxel_of(unsigned id, int nx, int ny, int nz, int& ix, int& iy, int& iz)
{
    if (id >= (unsigned)nx * (unsigned)ny * (unsigned)nz) return false;
    const unsigned r = id / (unsigned)nx;
    ix = (int)(id - r * (unsigned)nx);
    iy = (int)(r % (unsigned)ny);
    iz = (int)(r / (unsigned)ny);
    return true;
}

// out[i] = T(i + m(i) + 1/2)            (DeformKernel, deform_tex.cu.cc:15-26)
__global__ __launch_bounds__(256) void k_tex_deform(const float* __restrict__ img, const float* __restrict__ mx,
                                                    const float* __restrict__ my, const float* __restrict__ mz,
                                                    int nx, int ny, int nz, float* __restrict__ out)
{
    const unsigned id = blockIdx.x * 256u + threadIdx.x;
    int ix, iy, iz;
    if (!voxel_of(id, nx, ny, nz, ix, iy, iz)) return;
    out[id] = tex3(img, nx, ny, nz, (float)ix + mx[id] + 0.5f, (float)iy + my[id] + 0.5f, (float)iz + mz[id] + 0.5f);
}

// fixed-point inverse of a displacement field (DeformInvertKernel,
// deform_tex.cu.cc:78-100): d = (x, y, z) from 0; each iteration
// x = -mx(i + d), then y = -my(i + d), then z = -mz(i + d), every sample using
// the components already updated
__global__ __launch_bounds__(256) void k_tex_invert(const float* __restrict__ mx, const float* __restrict__ my,
                                                    const float* __restrict__ mz, int nx, int ny, int nz, int iters,
                                                    float* __restrict__ ox, float* __restrict__ oy,
                                                    float* __restrict__ oz)
{
    const unsigned id = blockIdx.x * 256u + threadIdx.x;
    int ix, iy, iz;
    if (!voxel_of(id, nx, ny, nz, ix, iy, iz)) return;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    const float fx = (float)ix, fy = (float)iy, fz = (float)iz;
    for (int it = 0; it < iters; ++it) {
        x = -tex3(mx, nx, ny, nz, x + fx + 0.5f, y + fy + 0.5f, z + fz + 0.5f);
        y = -tex3(my, nx, ny, nz, x + fx + 0.5f, y + fy + 0.5f, z + fz + 0.5f);
        z = -tex3(mz, nx, ny, nz, x + fx + 0.5f, y + fy + 0.5f, z + fz + 0.5f);
    }
    ox[id] = x;
    oy[id] = y;
    oz[id] = z;
}

// out[i] = T(i - s + 1/2)   (ShiftKernel :228-239; ShiftZKernel :291-302 with s = (0, 0, sz))
__global__ __launch_bounds__(256) void k_tex_shift(const float* __restrict__ img, int nx, int ny, int nz, float sx,
                                                   float sy, float sz, float* __restrict__ out)
{
    const unsigned id = blockIdx.x * 256u + threadIdx.x;
    int ix, iy, iz;
    if (!voxel_of(id, nx, ny, nz, ix, iy, iz)) return;
    out[id] = tex3(img, nx, ny, nz, (float)ix - sx + 0.5f, (float)iy - sy + 0.5f, (float)iz - sz + 0.5f);
}

// in-plane rotation about the integer centre (nx/2, ny/2) (RotateKernel :355-370)
__global__ __launch_bounds__(256) void k_tex_rotate(const float* __restrict__ img, int nx, int ny, int nz, float c,
                                                    float s, float* __restrict__ out)
{
    const unsigned id = blockIdx.x * 256u + threadIdx.x;
    int ix, iy, iz;
    if (!voxel_of(id, nx, ny, nz, ix, iy, iz)) return;
    const float hx = (float)(nx / 2), hy = (float)(ny / 2);
    const float ax = (float)ix + 0.5f - hx, ay = (float)iy + 0.5f - hy;
    const float xp = ax * c - ay * s + hx;
    const float yp = ax * s + ay * c + hy;
    out[id] = tex3(img, nx, ny, nz, xp, yp, (float)iz + 0.5f);
}

int check_grid(const int grid[3])
{
    if (!grid) return NEFCT_ERR_NULL;
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0) return NEFCT_ERR_SHAPE;
    const size_t n = (size_t)grid[0] * grid[1] * grid[2];
    if (n * 4 >= 0x7fffffffull) return NEFCT_ERR_SHAPE;
    return NEFCT_OK;
}

dim3 blocks_of(const int grid[3])
{
    const size_t n = (size_t)grid[0] * grid[1] * grid[2];
    return dim3((unsigned)((n + 255) / 256));
}

}  // namespace

extern "C" {

int nefct_deform_tex(const float* image, const float* mx, const float* my, const float* mz, const int grid[3],
                     float* deformed_image, void* stream)
{
    const int st = check_grid(grid);
    if (st) return st;
    if (!image || !mx || !my || !mz || !deformed_image) return NEFCT_ERR_NULL;
    hipLaunchKernelGGL(k_tex_deform, blocks_of(grid), dim3(256), 0, (hipStream_t)stream, image, mx, my, mz, grid[0],
                       grid[1], grid[2], deformed_image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_invert_dvf(const float* mx, const float* my, const float* mz, const int grid[3], int iterations,
                     float* inv_mx, float* inv_my, float* inv_mz, void* stream)
{
    const int st = check_grid(grid);
    if (st) return st;
    if (!mx || !my || !mz || !inv_mx || !inv_my || !inv_mz) return NEFCT_ERR_NULL;
    if (iterations < 0 || iterations > 10000) return NEFCT_ERR_SHAPE;
    hipLaunchKernelGGL(k_tex_invert, blocks_of(grid), dim3(256), 0, (hipStream_t)stream, mx, my, mz, grid[0],
                       grid[1], grid[2], iterations, inv_mx, inv_my, inv_mz);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_deform_invert_tex(const float* image, const float* mx, const float* my, const float* mz, const int grid[3],
                            float* deformed_image, void* workspace, size_t workspace_bytes, void* stream)
{
    const int st = check_grid(grid);
    if (st) return st;
    if (!image || !mx || !my || !mz || !deformed_image) return NEFCT_ERR_NULL;
    const size_t n = (size_t)grid[0] * grid[1] * grid[2];
    if (!workspace || workspace_bytes < 3 * n * sizeof(float)) return NEFCT_ERR_WORKSPACE;
    float* w = (float*)workspace;
    int s2 = nefct_invert_dvf(mx, my, mz, grid, 30, w, w + n, w + 2 * n, stream);
    if (s2) return s2;
    return nefct_deform_tex(image, w, w + n, w + 2 * n, grid, deformed_image, stream);
}

int nefct_shift_tex(const float* image, const int grid[3], float mx, float my, float mz, float* deformed_image,
                    void* stream)
{
    const int st = check_grid(grid);
    if (st) return st;
    if (!image || !deformed_image) return NEFCT_ERR_NULL;
    hipLaunchKernelGGL(k_tex_shift, blocks_of(grid), dim3(256), 0, (hipStream_t)stream, image, grid[0], grid[1],
                       grid[2], mx, my, mz, deformed_image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_shift_z_tex(const float* image, const int grid[3], float mz, float* deformed_image, void* stream)
{
    return nefct_shift_tex(image, grid, 0.0f, 0.0f, mz, deformed_image, stream);
}

int nefct_rotate_tex(const float* image, const int grid[3], float cos_phi, float sin_phi, float* deformed_image,
                     void* stream)
{
    const int st = check_grid(grid);
    if (st) return st;
    if (!image || !deformed_image) return NEFCT_ERR_NULL;
    hipLaunchKernelGGL(k_tex_rotate, blocks_of(grid), dim3(256), 0, (hipStream_t)stream, image, grid[0], grid[1],
                       grid[2], cos_phi, sin_phi, deformed_image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // extern "C"
