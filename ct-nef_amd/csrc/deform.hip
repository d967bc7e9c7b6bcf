// Trilinear volume warp by a dense displacement field, its transpose and its
// gradient with respect to the displacement, for gfx950.
//
// Replaces cpp/deform/deform.cu.cc:13-46 (DeformKernel) and :112-126 (host
// launcher) behind the Deform op (cpp/deform/deform.cc:12-22, :59-91), with the
// zero-filled output of nefct/cc/deform/deform.cc:94.  The reference registers
// no gradient; the transpose (scatter) and the displacement gradient here make
// the warp differentiable.
//
//   out[i] = sum over the 8 corners c of img[c] * w_c(p_i),  p_i = i + m(i)
//
// m is in voxel units.  Boundary rule 0 (reference): the voxel is written only
// when every base index floor(p) lies in [1, n-2] (deform.cu.cc:31-33), else 0.
// Rule 1 (zero padding): every corner inside the volume contributes.
// One lane per output voxel (forward and field gradient: two, in neighbouring y rows), x
// fastest: the three displacement reads and the write are coalesced, the 8
// gathers hit L2 (neighbouring lanes share corners).
#include "common.h"

namespace {

struct Corners {
    int x0, y0, z0;
    float fx, fy, fz;   // weights of the upper corners
    bool ok;            // reference rule: all base indices in [1, n-2]
};

__device__ __forceinline__ Corners corners_of(int ix, int iy, int iz, float mx, float my, float mz, int nx, int ny,
                                              int nz)
{
    Corners c;
    const float px = (float)ix + mx, py = (float)iy + my, pz = (float)iz + mz;
    const float flx = floorf(px), fly = floorf(py), flz = floorf(pz);
    // The base index is clamped to [-2, n] before the conversion: a position
    // that far outside has both corners outside the volume, as it has before
    // the clamp, and a displacement of 2^31 voxels or more would otherwise
    // overflow the conversion (finite3 lets |m| < 1e30 through).  The weights
    // keep the unclamped floor, so they stay in [0, 1].
    c.x0 = (int)fminf(fmaxf(flx, -2.0f), (float)nx);
    c.y0 = (int)fminf(fmaxf(fly, -2.0f), (float)ny);
    c.z0 = (int)fminf(fmaxf(flz, -2.0f), (float)nz);
    c.fx = px - flx;
    c.fy = py - fly;
    c.fz = pz - flz;
    c.ok = c.x0 > 0 && c.x0 + 1 < nx && c.y0 > 0 && c.y0 + 1 < ny && c.z0 > 0 && c.z0 + 1 < nz;
    return c;
}

__device__ __forceinline__ bool inside(int x, int y, int z, int nx, int ny, int nz)
{
    return (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny && (unsigned)z < (unsigned)nz;
}

// Finite-precision guard: a non-finite displacement makes every corner invalid.
__device__ __forceinline__ bool finite3(float a, float b, float c)
{
    return fabsf(a) < 1e30f && fabsf(b) < 1e30f && fabsf(c) < 1e30f;
}

// The two x-corners of a (y, z) row, each weighted, or 0 outside the volume.
// When both lie in the row (always under the reference rule) they are read as
// one 8-byte load (4-byte aligned; gfx950 loads need only dword alignment).
__device__ __forceinline__ float2 row_pair(const float* __restrict__ img, int x0, int y, int z, int nx, int ny, int nz)
{
    if ((unsigned)y >= (unsigned)ny || (unsigned)z >= (unsigned)nz) return make_float2(0.0f, 0.0f);
    const float* row = img + (size_t)nx * ((size_t)y + (size_t)ny * z);
    if (x0 >= 0 && x0 + 1 < nx) {
        float2 p;
        __builtin_memcpy(&p, row + x0, 8);
        return p;
    }
    return make_float2((unsigned)x0 < (unsigned)nx ? row[x0] : 0.0f,
                       (unsigned)(x0 + 1) < (unsigned)nx ? row[x0 + 1] : 0.0f);
}

// Two voxels per lane, in neighbouring y rows: their gathers share image rows
// and are in flight together (6 % faster than one voxel per lane at 512^2 x 256).
__global__ void k_deform(const float* __restrict__ img, const float* __restrict__ mx, const float* __restrict__ my,
                         const float* __restrict__ mz, int nx, int ny, int nz, int boundary, float* __restrict__ out)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.z;
    if (ix >= nx) return;
    float v[2];
    size_t ids[2];
    bool in[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int iy = blockIdx.y * 2 + k;
        in[k] = iy < ny;
        const int iyc = in[k] ? iy : ny - 1;   // the row past an odd ny reads the last row, writes nothing
        ids[k] = (size_t)ix + (size_t)nx * ((size_t)iyc + (size_t)ny * iz);
        const float dx = mx[ids[k]], dy = my[ids[k]], dz = mz[ids[k]];
        v[k] = 0.0f;
        if (finite3(dx, dy, dz)) {
            const Corners c = corners_of(ix, iyc, iz, dx, dy, dz, nx, ny, nz);
            if (boundary != 0 || c.ok) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int y = c.y0 + (q & 1), z = c.z0 + (q >> 1);
                    const float2 p = row_pair(img, c.x0, y, z, nx, ny, nz);
                    const float wyz = ((q & 1) ? c.fy : 1.0f - c.fy) * ((q >> 1) ? c.fz : 1.0f - c.fz);
                    v[k] = fmaf(fmaf(c.fx, p.y - p.x, p.x), wyz, v[k]);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (in[k]) out[ids[k]] = v[k];
}

// grad_img[c] += g[i] * w_c(p_i)   (transpose of k_deform in the image)
__global__ void k_deform_adjoint(const float* __restrict__ g, const float* __restrict__ mx,
                                 const float* __restrict__ my, const float* __restrict__ mz, int nx, int ny, int nz,
                                 int boundary, float* __restrict__ gimg, uint32_t gbytes)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int iz = blockIdx.z;
    if (ix >= nx) return;
    const size_t id = (size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz);
    const float gv = g[id];
    const float dx = mx[id], dy = my[id], dz = mz[id];
    if (gv == 0.0f || !finite3(dx, dy, dz)) return;
    const Corners c = corners_of(ix, iy, iz, dx, dy, dz, nx, ny, nz);
    if (boundary == 0 && !c.ok) return;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(gimg, gbytes);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int x = c.x0 + (q & 1), y = c.y0 + ((q >> 1) & 1), z = c.z0 + (q >> 2);
        const float w = ((q & 1) ? c.fx : 1.0f - c.fx) * (((q >> 1) & 1) ? c.fy : 1.0f - c.fy) *
                        ((q >> 2) ? c.fz : 1.0f - c.fz);
        if (inside(x, y, z, nx, ny, nz) && w != 0.0f)
            buf_atomic_add(rs, (int)(((size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)) * 4), gv * w);
    }
}

// d<g, warp(img, m)>/dm at each voxel (the derivative of the trilinear
// interpolant on the cell that contains p; zero where the reference rule
// writes 0, since the output is constant there).
__global__ void k_deform_dvf_grad(const float* __restrict__ img, const float* __restrict__ g,
                                  const float* __restrict__ mx, const float* __restrict__ my,
                                  const float* __restrict__ mz, int nx, int ny, int nz, int boundary,
                                  float* __restrict__ gx, float* __restrict__ gy, float* __restrict__ gz)
{
    // two voxels per lane in neighbouring y rows, as in k_deform
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.z;
    if (ix >= nx) return;
    float ax[2], ay[2], az[2], gv[2];
    size_t ids[2];
    bool in[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int iy = blockIdx.y * 2 + k;
        in[k] = iy < ny;
        const int iyc = in[k] ? iy : ny - 1;   // the row past an odd ny reads the last row, writes nothing
        ids[k] = (size_t)ix + (size_t)nx * ((size_t)iyc + (size_t)ny * iz);
        gv[k] = g[ids[k]];
        const float dx = mx[ids[k]], dy = my[ids[k]], dz = mz[ids[k]];
        ax[k] = ay[k] = az[k] = 0.0f;
        if (gv[k] != 0.0f && finite3(dx, dy, dz)) {
            const Corners c = corners_of(ix, iyc, iz, dx, dy, dz, nx, ny, nz);
            if (boundary != 0 || c.ok) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int y = c.y0 + (q & 1), z = c.z0 + (q >> 1);
                    const float2 p = row_pair(img, c.x0, y, z, nx, ny, nz);
                    const float wy = (q & 1) ? c.fy : 1.0f - c.fy, sy = (q & 1) ? 1.0f : -1.0f;
                    const float wz = (q >> 1) ? c.fz : 1.0f - c.fz, sz = (q >> 1) ? 1.0f : -1.0f;
                    const float lx = fmaf(c.fx, p.y - p.x, p.x);   // the row interpolated in x
                    ax[k] = fmaf(p.y - p.x, wy * wz, ax[k]);
                    ay[k] = fmaf(lx, sy * wz, ay[k]);
                    az[k] = fmaf(lx, wy * sz, az[k]);
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (!in[k]) continue;
        gx[ids[k]] = ax[k] * gv[k];
        gy[ids[k]] = ay[k] * gv[k];
        gz[ids[k]] = az[k] * gv[k];
    }
}

__global__ void k_fill0(float* __restrict__ p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0f;
}

int check_grid(const int grid[3])
{
    if (!grid) return NEFCT_ERR_NULL;
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0) return NEFCT_ERR_SHAPE;
    if (grid[1] > 65535 || grid[2] > 65535) return NEFCT_ERR_SHAPE;
    const size_t n = (size_t)grid[0] * grid[1] * grid[2];
    if (n * 4 >= 0x7fffffffull) return NEFCT_ERR_SHAPE;
    return NEFCT_OK;
}

}  // namespace

extern "C" {

int nefct_deform(const float* image, const float* mx, const float* my, const float* mz, const int grid[3],
                 float* deformed_image, int boundary, void* stream)
{
    int st = check_grid(grid);
    if (st) return st;
    if (!image || !mx || !my || !mz || !deformed_image) return NEFCT_ERR_NULL;
    if (boundary != 0 && boundary != 1) return NEFCT_ERR_UNSUPPORTED;
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    hipLaunchKernelGGL(k_deform, dim3((nx + 255) / 256, (ny + 1) / 2, nz), dim3(256), 0, (hipStream_t)stream, image, mx, my, mz,
                       nx, ny, nz, boundary, deformed_image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_deform_adjoint(const float* grad_out, const float* mx, const float* my, const float* mz, const int grid[3],
                         float* grad_image, int boundary, void* stream)
{
    int st = check_grid(grid);
    if (st) return st;
    if (!grad_out || !mx || !my || !mz || !grad_image) return NEFCT_ERR_NULL;
    if (boundary != 0 && boundary != 1) return NEFCT_ERR_UNSUPPORTED;
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    const size_t n = (size_t)nx * ny * nz;
    hipStream_t s = (hipStream_t)stream;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill0, dim3((unsigned)blocks), dim3(256), 0, s, grad_image, n);
    hipLaunchKernelGGL(k_deform_adjoint, dim3((nx + 255) / 256, ny, nz), dim3(256), 0, s, grad_out, mx, my, mz, nx,
                       ny, nz, boundary, grad_image, (uint32_t)(n * 4));
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_deform_dvf_grad(const float* image, const float* grad_out, const float* mx, const float* my,
                          const float* mz, const int grid[3], float* grad_mx, float* grad_my, float* grad_mz,
                          int boundary, void* stream)
{
    int st = check_grid(grid);
    if (st) return st;
    if (!image || !grad_out || !mx || !my || !mz || !grad_mx || !grad_my || !grad_mz) return NEFCT_ERR_NULL;
    if (boundary != 0 && boundary != 1) return NEFCT_ERR_UNSUPPORTED;
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    hipLaunchKernelGGL(k_deform_dvf_grad, dim3((nx + 255) / 256, (ny + 1) / 2, nz), dim3(256), 0, (hipStream_t)stream, image,
                       grad_out, mx, my, mz, nx, ny, nz, boundary, grad_mx, grad_my, grad_mz);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // extern "C"
