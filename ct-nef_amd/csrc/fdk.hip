// FDK reconstruction stages for gfx950 (flat detector): pre-weighting, ramp
// filtering and voxel-driven weighted backprojection.
//
// Replaces the FDK path of the reference, which exists as
//   * MATLAB weighting and filtering: cosine weight mat_codes/filtering.m:5-14,
//     Parker short-scan weights mat_codes/ParkerWeight.m:5-27, ramp filter with
//     windows and scaling mat_codes/filtering.m:20-48,54-88;
//   * a detector-driven CUDA backprojection with the (SOD/U)^2 weight and
//     helical per-view z offsets, nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-298
//     (no op registration), and the voxel-driven MATLAB backprojection
//     mat_codes/backprojection.m:11-25 whose bilinear sampling is used here.
//
// Geometry is the siddon-op convention (SID = source-isocentre, SAD =
// source-detector, a = (ai + ia) da - na da / 2); view offsets shift source and
// detector along z as in dist_3d_back_fdk.cu.cc:46,65.
#include <math.h>

#include "common.h"

namespace {

// q[iv][ib][ia] = p * SDD / sqrt(SDD^2 + a^2 + b^2) * parker(iv, ia) * scale
__global__ void k_fdk_preweight(const float* __restrict__ p, float* __restrict__ q, int nv, int na, int nb,
                                double SAD, double da, double ai, double db, double bi, int parker, double dang_deg,
                                int direction, double scale)
{
    const int ia = blockIdx.x * blockDim.x + threadIdx.x;
    const int ib = blockIdx.y;
    const int iv = blockIdx.z;
    if (ia >= na) return;
    const double a = (ai + ia) * da - na * da * 0.5;
    const double b = (bi + ib) * db - nb * db * 0.5;
    double w = SAD / sqrt(SAD * SAD + a * a + b * b) * scale;
    if (parker) {
        // fan angles of the first and last column and of this one, degrees
        const double rad2deg = 57.295779513082320876798;
        const double a0 = (ai + 0) * da - na * da * 0.5, a1 = (ai + (na - 1)) * da - na * da * 0.5;
        const double g0 = atan(a0 / SAD) * rad2deg, g1 = atan(a1 / SAD) * rad2deg;
        const double gam = atan(a / SAD) * rad2deg;
        const double fan = fabs(g1 - g0) * 0.5;
        const double n180 = 180.0 / dang_deg;
        const double pw = ((n180 - nv) * 0.5 + (iv + 1)) / n180 * 180.0 + direction * gam;
        const double pi2 = 1.5707963267948966192313;
        double pk = 1.0;
        if (pw <= fan * 0.5 && pw > -fan * 0.5) {
            const double sv = sin(pi2 * (fan * 0.5 + pw) / fan);
            pk = sv > 0.0 ? sv * sv : 0.0;
        } else if (pw >= 180.0 - fan * 0.5 && pw < 180.0 + fan * 0.5) {
            const double cv = cos(pi2 * (fan * 0.5 + pw - 180.0) / fan);
            pk = cv > 0.0 ? cv * cv : 0.0;
        } else if (pw < -fan * 0.5 || pw > 180.0 + fan * 0.5) {
            pk = 0.0;
        }
        w *= pk;
    }
    const size_t id = (size_t)ia + (size_t)na * ((size_t)ib + (size_t)nb * iv);
    q[id] = (float)(p[id] * w);
}

// Row convolution with a (2 na - 1)-tap kernel h centred at na - 1:
//   out[u] = sum_v in[v] * h[u - v + na - 1]
// One workgroup per FR rows; rows and taps staged in LDS; every thread
// produces 4 neighbouring outputs of one row with a sliding tap window.
constexpr int FR = 4;
constexpr int FMAX = 2048;   // longest detector row supported

__global__ __launch_bounds__(256) void k_fdk_filter(const float* __restrict__ in, float* __restrict__ out,
                                                    const float* __restrict__ h, int rows, int na)
{
    __shared__ float srow[FR][FMAX];
    __shared__ float sh[2 * FMAX];
    const int r0 = blockIdx.x * FR;
    const int tid = threadIdx.x;
    const int ntap = 2 * na - 1;
    for (int i = tid; i < ntap; i += 256) sh[i] = h[i];
    sh[ntap] = 0.0f;
    for (int r = 0; r < FR; ++r) {
        const int row = r0 + r;
        for (int i = tid; i < na; i += 256) srow[r][i] = (row < rows) ? in[(size_t)row * na + i] : 0.0f;
    }
    __syncthreads();
    const int nq = (na + 3) / 4;  // groups of 4 outputs per row
    for (int job = tid; job < FR * nq; job += 256) {
        const int r = job / nq, u0 = (job % nq) * 4;
        if (r0 + r >= rows) continue;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
        // tap index for output u and input v: u - v + na - 1; start at v = 0
        int k = u0 + na - 1;
        float h0 = sh[k], h1 = (k + 1 < ntap) ? sh[k + 1] : 0.0f, h2 = (k + 2 < ntap) ? sh[k + 2] : 0.0f,
              h3 = (k + 3 < ntap) ? sh[k + 3] : 0.0f;
        for (int v = 0; v < na; ++v) {
            const float x = srow[r][v];
            a0 = fmaf(x, h0, a0);
            a1 = fmaf(x, h1, a1);
            a2 = fmaf(x, h2, a2);
            a3 = fmaf(x, h3, a3);
            // next input: taps shift down by one
            h3 = h2;
            h2 = h1;
            h1 = h0;
            --k;
            h0 = (k >= 0) ? sh[k] : 0.0f;
        }
        const size_t base = (size_t)(r0 + r) * na;
        if (u0 + 0 < na) out[base + u0 + 0] = a0;
        if (u0 + 1 < na) out[base + u0 + 1] = a1;
        if (u0 + 2 < na) out[base + u0 + 2] = a2;
        if (u0 + 3 < na) out[base + u0 + 3] = a3;
    }
}

// Voxel-driven weighted backprojection.  Thread = (ix, iy, z-chunk of ZC
// voxels); all views accumulate in registers; one write per voxel.
constexpr int ZC = 8;

__global__ __launch_bounds__(256) void k_fdk_bp(const float* __restrict__ q, const double* __restrict__ trig,
                                                const float* __restrict__ offsets, int nv, int nx, int ny, int nz,
                                                double x0, double y0, double z0, double vx, double vy, double vz,
                                                double SID, double SAD, int na, double da, double ai, int nb,
                                                double db, double bi, float* __restrict__ image)
{
    const int ix = blockIdx.x * 64 + (threadIdx.x & 63);
    const int iy = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int izc = blockIdx.z * ZC;
    if (ix >= nx || iy >= ny) return;
    const double x = x0 + (ix + 0.5) * vx, y = y0 + (iy + 0.5) * vy;
    float acc[ZC];
#pragma unroll
    for (int k = 0; k < ZC; ++k) acc[k] = 0.0f;
    const double ha = na * da * 0.5, hb = nb * db * 0.5;
    const float inv_db = (float)(1.0 / db);
    for (int iv = 0; iv < nv; ++iv) {
        const double c = trig[2 * iv], s = trig[2 * iv + 1];
        const double U = SID + x * c + y * s;
        if (!(U > 0.0)) continue;
        const double mag = SAD / U;
        const double fa = ((-x * s + y * c) * mag + ha) / da - ai;
        if (!(fa >= 0.0 && fa <= (double)(na - 1))) continue;
        const int ia0 = min((int)floor(fa), max(na - 2, 0));
        const float ta = (float)(fa - ia0);
        const float w = (float)((SID / U) * (SID / U));
        const float off = offsets ? offsets[iv] : 0.0f;
        const float magf = (float)mag;
        const float fb0 = (float)((hb) / db - bi);
        const float* qv = q + (size_t)nb * na * iv;
        const int ia1 = min(ia0 + 1, na - 1);
#pragma unroll
        for (int k = 0; k < ZC; ++k) {
            const int iz = izc + k;
            if (iz >= nz) break;
            const float zr = (float)(z0 + (iz + 0.5) * vz) - off;
            const float fb = fmaf(zr * magf, inv_db, fb0);
            if (!(fb >= 0.0f && fb <= (float)(nb - 1))) continue;
            const int ib0 = min((int)floorf(fb), max(nb - 2, 0));
            const int ib1 = min(ib0 + 1, nb - 1);
            const float tb = fb - (float)ib0;
            const float* r0p = qv + (size_t)ib0 * na;
            const float* r1p = qv + (size_t)ib1 * na;
            const float top = fmaf(ta, r0p[ia1] - r0p[ia0], r0p[ia0]);
            const float bot = fmaf(ta, r1p[ia1] - r1p[ia0], r1p[ia0]);
            acc[k] = fmaf(w, fmaf(tb, bot - top, top), acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < ZC; ++k) {
        const int iz = izc + k;
        if (iz < nz) image[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)] = acc[k];
    }
}

__global__ void k_trig(const float* __restrict__ angles, int nv, double* __restrict__ trig)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const double a = (double)angles[i];
    trig[2 * i] = cos(a);
    trig[2 * i + 1] = sin(a);
}

bool finite_pos(double v) { return std::isfinite(v) && v > 0.0; }

}  // namespace

extern "C" {

int nefct_fdk_preweight(const float* projection_value, float* out, int nv, float SID, float SAD, int na, float da,
                        float ai, int nb, float db, float bi, int parker, float dang_deg, int direction, float scale,
                        void* stream)
{
    if (!projection_value || !out) return NEFCT_ERR_NULL;
    if (nv <= 0 || na <= 0 || nb <= 0 || nb > 65535 || nv > 65535) return NEFCT_ERR_SHAPE;
    if (!finite_pos(SID) || !finite_pos(SAD) || !finite_pos(da) || !finite_pos(db) || !std::isfinite(ai) ||
        !std::isfinite(bi) || !std::isfinite(scale))
        return NEFCT_ERR_GEOMETRY;
    if (parker && (!finite_pos(dang_deg) || (direction != 1 && direction != -1))) return NEFCT_ERR_GEOMETRY;
    hipLaunchKernelGGL(k_fdk_preweight, dim3((na + 255) / 256, nb, nv), dim3(256), 0, (hipStream_t)stream,
                       projection_value, out, nv, na, nb, (double)SAD, (double)da, (double)ai, (double)db,
                       (double)bi, parker, (double)dang_deg, direction, (double)scale);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_fdk_filter(const float* in, float* out, int rows, int na, const float* taps, void* stream)
{
    if (!in || !out || !taps) return NEFCT_ERR_NULL;
    if (rows <= 0 || na <= 0 || na > FMAX) return NEFCT_ERR_SHAPE;
    if (in == out) return NEFCT_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(k_fdk_filter, dim3((rows + FR - 1) / FR), dim3(256), 0, (hipStream_t)stream, in, out, taps,
                       rows, na);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

size_t nefct_fdk_workspace_size(int nv) { return sizeof(double) * 2 * (size_t)(nv > 0 ? nv : 1); }

int nefct_fdk_backproject(const float* filtered, const int grid[3], const float center[3], const float size[3],
                          const float* angles, const float* offsets, int nv, float SID, float SAD, int na, float da,
                          float ai, int nb, float db, float bi, float* image, void* workspace,
                          size_t workspace_bytes, void* stream)
{
    if (!filtered || !grid || !center || !size || !angles || !image) return NEFCT_ERR_NULL;
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0 || nv <= 0 || na <= 0 || nb <= 0) return NEFCT_ERR_SHAPE;
    if ((grid[1] + 3) / 4 > 65535 || (grid[2] + ZC - 1) / ZC > 65535) return NEFCT_ERR_SHAPE;
    if (!finite_pos(SID) || !finite_pos(SAD) || !finite_pos(da) || !finite_pos(db) || !finite_pos(size[0]) ||
        !finite_pos(size[1]) || !finite_pos(size[2]) || !std::isfinite(center[0]) || !std::isfinite(center[1]) ||
        !std::isfinite(center[2]) || !std::isfinite(ai) || !std::isfinite(bi))
        return NEFCT_ERR_GEOMETRY;
    if (!workspace || workspace_bytes < nefct_fdk_workspace_size(nv)) return NEFCT_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    double* trig = (double*)workspace;
    hipLaunchKernelGGL(k_trig, dim3((nv + 255) / 256), dim3(256), 0, s, angles, nv, trig);
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    const double vx = (double)size[0] / nx, vy = (double)size[1] / ny, vz = (double)size[2] / nz;
    const double x0 = (double)center[0] - 0.5 * size[0], y0 = (double)center[1] - 0.5 * size[1],
                 z0 = (double)center[2] - 0.5 * size[2];
    hipLaunchKernelGGL(k_fdk_bp, dim3((nx + 63) / 64, (ny + 3) / 4, (nz + ZC - 1) / ZC), dim3(256), 0, s, filtered,
                       (const double*)trig, offsets, nv, nx, ny, nz, x0, y0, z0, vx, vy, vz, (double)SID,
                       (double)SAD, na, (double)da, (double)ai, nb, (double)db, (double)bi, image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // extern "C"
