// Smoothed anisotropic total variation for gfx950: value, gradient and the
// fused gradient-descent update of the iterative reconstruction.
//
// TV(x) = sum over voxels and axes k of sqrt(d_k^2 + eps^2), d_k the forward
// difference along k with a zero first slice (the differences of
// nefct/functions/total_variation.py:29-35; their transpose is
// InvertTotalVari, :125-149).  d TV / d x_j = sum_k phi(d_k(j)) - phi(d_k(j + e_k)),
// phi(d) = d / sqrt(d^2 + eps^2), where the terms whose difference does not
// exist (first slice, past the last) are dropped.
//
// One lane per voxel, a wavefront = 64 consecutive x; the 7-point stencil is
// read straight from HBM/L2 (the y and z neighbours were read by the lanes of
// the neighbouring rows/planes shortly before, so they hit L2).  The value is
// reduced per workgroup and added to one FP64 accumulator.
#include <math.h>

#include <cmath>

#include "common.h"

namespace {

__device__ __forceinline__ float phi(float d, float e2) { return d * rsqrtf(d * d + e2); }

__device__ __forceinline__ float tv_grad_at(const float* __restrict__ x, int ix, int iy, int iz, int nx, int ny,
                                            int nz, float e2, float& val)
{
    const size_t sy = nx, sz = (size_t)nx * ny;
    const size_t o = ix + sy * iy + sz * iz;
    const float c = x[o];
    float g = 0.0f;
    val = 0.0f;
    // own differences (backward neighbour): contribute to value and +phi
    if (ix > 0) { const float d = c - x[o - 1]; g += phi(d, e2); val += sqrtf(d * d + e2); }
    else val += sqrtf(e2);
    if (iy > 0) { const float d = c - x[o - sy]; g += phi(d, e2); val += sqrtf(d * d + e2); }
    else val += sqrtf(e2);
    if (iz > 0) { const float d = c - x[o - sz]; g += phi(d, e2); val += sqrtf(d * d + e2); }
    else val += sqrtf(e2);
    // forward neighbours' differences: -phi
    if (ix + 1 < nx) g -= phi(x[o + 1] - c, e2);
    if (iy + 1 < ny) g -= phi(x[o + sy] - c, e2);
    if (iz + 1 < nz) g -= phi(x[o + sz] - c, e2);
    return g;
}

// grad (may be null) = scale * dTV/dx; *value (may be null) += scale * TV(x)
__global__ __launch_bounds__(256) void k_tv(const float* __restrict__ x, int nx, int ny, int nz, float e2,
                                            float scale, float* __restrict__ grad, double* __restrict__ value)
{
    __shared__ double red[4];
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    double v = 0.0;
    if (ix < nx && iy < ny) {
        float val;
        const float g = tv_grad_at(x, ix, iy, iz, nx, ny, nz, e2, val);
        if (grad) grad[ix + (size_t)nx * (iy + (size_t)ny * iz)] = scale * g;
        v = val;
    }
    if (!value) return;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(value, (double)scale * (red[0] + red[1] + red[2] + red[3]));
}

// *value += scale * sum x^2 (grid-stride; FP64 partial sums)
__global__ __launch_bounds__(256) void k_sqnorm(const float* __restrict__ x, size_t n, double scale,
                                                double* __restrict__ value)
{
    __shared__ double red[4];
    double v = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double t = x[i];
        v += t * t;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(value, scale * (red[0] + red[1] + red[2] + red[3]));
}

// out = x - step * (data_grad + lam * dTV/dx)   (out != x)
__global__ __launch_bounds__(256) void k_tv_descent(const float* __restrict__ x, const float* __restrict__ data_grad,
                                                    int nx, int ny, int nz, float e2, float lam, float step,
                                                    float* __restrict__ out)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int iz = blockIdx.z;
    if (ix >= nx || iy >= ny) return;
    float val;
    const float g = tv_grad_at(x, ix, iy, iz, nx, ny, nz, e2, val);
    const size_t o = ix + (size_t)nx * (iy + (size_t)ny * iz);
    out[o] = x[o] - step * (data_grad[o] + lam * g);
}

int check_grid(const int grid[3])
{
    if (!grid) return NEFCT_ERR_NULL;
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0) return NEFCT_ERR_SHAPE;
    if ((grid[1] + 3) / 4 > 65535 || grid[2] > 65535) return NEFCT_ERR_SHAPE;
    return NEFCT_OK;
}

}  // namespace

extern "C" {

int nefct_tv_smooth(const float* image, const int grid[3], float eps, float scale, float* grad, double* value,
                    void* stream)
{
    int st = check_grid(grid);
    if (st) return st;
    if (!image || (!grad && !value)) return NEFCT_ERR_NULL;
    if (!(eps > 0.0f) || !std::isfinite(eps) || !std::isfinite(scale)) return NEFCT_ERR_GEOMETRY;
    const dim3 g((grid[0] + 63) / 64, (grid[1] + 3) / 4, grid[2]);
    hipLaunchKernelGGL(k_tv, g, dim3(256), 0, (hipStream_t)stream, image, grid[0], grid[1], grid[2], eps * eps,
                       scale, grad, value);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_sq_norm(const float* x, size_t n, double scale, double* value, void* stream)
{
    if (!x || !value) return NEFCT_ERR_NULL;
    if (!std::isfinite(scale)) return NEFCT_ERR_GEOMETRY;
    if (n == 0) return NEFCT_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_sqnorm, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, x, n, scale, value);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_tv_descent_step(const float* image, const float* data_grad, const int grid[3], float eps, float lam,
                          float step, float* out, void* stream)
{
    int st = check_grid(grid);
    if (st) return st;
    if (!image || !data_grad || !out) return NEFCT_ERR_NULL;
    if (out == image) return NEFCT_ERR_UNSUPPORTED;
    if (!(eps > 0.0f) || !std::isfinite(eps) || !std::isfinite(lam) || !std::isfinite(step))
        return NEFCT_ERR_GEOMETRY;
    const dim3 g((grid[0] + 63) / 64, (grid[1] + 3) / 4, grid[2]);
    hipLaunchKernelGGL(k_tv_descent, g, dim3(256), 0, (hipStream_t)stream, image, data_grad, grid[0], grid[1],
                       grid[2], eps * eps, lam, step, out);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // extern "C"
