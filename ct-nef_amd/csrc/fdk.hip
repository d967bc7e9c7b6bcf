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
#include "fdk_staged.h"

namespace {

// q[iv][ib][ia] = p * SDD / sqrt(SDD^2 + a^2 + b^2) * parker(iv, ia) * scale
// Thread = (column ia, view iv) over a run of PW_ROWS detector rows: the
// Parker factor (a function of the view and the column) is computed once.
constexpr int PW_ROWS = 64;

__global__ __launch_bounds__(256) void k_fdk_preweight(const float* __restrict__ p, float* __restrict__ q, int nv,
                                                       int na, int nb, double SAD, double da, double ai, double db,
                                                       double bi, int parker, double dang_deg, int direction,
                                                       double scale)
{
    const int ia = blockIdx.x * 256 + threadIdx.x;
    const int iv = blockIdx.y;
    const int ib0 = blockIdx.z * PW_ROWS, ib1 = min(nb, ib0 + PW_ROWS);
    if (ia >= na) return;
    const double a = (ai + ia) * da - na * da * 0.5;
    double pk = 1.0;
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
        if (pw <= fan * 0.5 && pw > -fan * 0.5) {
            const double sv = sin(pi2 * (fan * 0.5 + pw) / fan);
            pk = sv > 0.0 ? sv * sv : 0.0;
        } else if (pw >= 180.0 - fan * 0.5 && pw < 180.0 + fan * 0.5) {
            const double cv = cos(pi2 * (fan * 0.5 + pw - 180.0) / fan);
            pk = cv > 0.0 ? cv * cv : 0.0;
        } else if (pw < -fan * 0.5 || pw > 180.0 + fan * 0.5) {
            pk = 0.0;
        }
    }
    const double sa = SAD * SAD + a * a;
    const double f = SAD * scale * pk;
    for (int ib = ib0; ib < ib1; ++ib) {
        const double b = (bi + ib) * db - nb * db * 0.5;
        const size_t id = (size_t)ia + (size_t)na * ((size_t)ib + (size_t)nb * iv);
        q[id] = (float)(p[id] * (f / sqrt(sa + b * b)));
    }
}

// Row convolution with a (2 na - 1)-tap kernel h centred at na - 1:
//   out[u] = sum_v in[v] * h[u - v + na - 1]
// computed as a circular convolution of length N (the smallest power of two
// >= 2 na - 1, at least 64): kernel offsets |u - v| <= na - 1 < N / 2, so
// the first na outputs of the circular convolution are exactly the linear
// ones (the reference filters the same way, FFT of a zero-padded row,
// filtering.m:29-41).  Each workgroup transforms the taps once into an LDS
// response, then takes two rows at a time as the real and imaginary parts of
// one complex row (the taps are real, so the two filtered rows come back as
// the real and imaginary parts): FFT, multiply, inverse FFT through
// conj(FFT(conj(.))) / N.  FFTs are radix-4 Stockham (self-sorting) in LDS.
constexpr int FMAX = 2048;   // longest detector row supported (N <= 4096)

// Forward DFT of buf[a] (length N = 2^k) by self-sorting Stockham passes:
// radix 8 while a factor of 8 remains, then one radix-4 or radix-2 pass
// (2048 = 8 8 8 4: four LDS round trips).  tw[t] = exp(-2 pi i t / N) for
// t < N / 2; exp(-2 pi i (t + N/2) / N) = -tw[t].  Returns the buffer that
// holds the result.
template <int N>
__device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int t)
{
    return t < N / 2 ? tw[t] : make_float2(-tw[t - N / 2].x, -tw[t - N / 2].y);
}

// With SRC the first pass (radix 8, Ns = 1) takes its inputs from
// src(u, m, j) = element j + m N/8 (j = threadIdx.x + 256 u) instead of
// buf[a]; with SNK the last pass hands each output (index, value) to snk
// instead of storing it.  Both save an LDS round trip and a barrier.
template <int N, bool SRC = false, bool SNK = false, class Src, class Snk>
__device__ __forceinline__ int fft_stockham(float2 (*buf)[N], int a, const float2* __restrict__ tw, Src&& src,
                                            Snk&& snk)
{
    auto cmul = [](float2 u, float2 w) { return make_float2(u.x * w.x - u.y * w.y, u.x * w.y + u.y * w.x); };
    auto add = [](float2 u, float2 v) { return make_float2(u.x + v.x, u.y + v.y); };
    auto sub = [](float2 u, float2 v) { return make_float2(u.x - v.x, u.y - v.y); };
    auto mi = [](float2 u) { return make_float2(u.y, -u.x); };   // u * (-i)
    constexpr float R2 = 0.70710678118654752f;
    int Ns = 1;
    for (; Ns * 8 <= N; Ns *= 8) {
        constexpr int E = N / 8;
        const int step = N / (Ns * 8);
        const bool first = Ns == 1, last = Ns * 8 == N;
#pragma unroll
        for (int u = 0; u < (E + 255) / 256; ++u) {
            const int j = threadIdx.x + 256 * u;
            if (j >= E) break;
            const int k = j & (Ns - 1);
            float2 v[8];
            if (SRC && first) {
#pragma unroll
                for (int m = 0; m < 8; ++m) v[m] = src(u, m, j);   // Ns = 1: every twiddle is 1
            } else {
                v[0] = buf[a][j];
#pragma unroll
                for (int m = 1; m < 8; ++m) v[m] = cmul(buf[a][j + m * E], twiddle<N>(tw, m * k * step));
            }
            // 8-point DFT, decimation in frequency: even outputs from
            // v_m + v_{m+4}, odd ones from (v_m - v_{m+4}) W8^m
            const float2 b0 = add(v[0], v[4]), b1 = add(v[1], v[5]), b2 = add(v[2], v[6]), b3 = add(v[3], v[7]);
            const float2 d0 = sub(v[0], v[4]);
            const float2 t1 = sub(v[1], v[5]), t3 = sub(v[3], v[7]);
            const float2 d1 = make_float2(R2 * (t1.x + t1.y), R2 * (t1.y - t1.x));     // W8^1
            const float2 d2 = mi(sub(v[2], v[6]));                                   // W8^2
            const float2 d3 = make_float2(R2 * (t3.y - t3.x), -R2 * (t3.x + t3.y));    // W8^3
            const float2 c0 = add(b0, b2), c2 = sub(b0, b2), c1 = add(b1, b3), c3 = mi(sub(b1, b3));
            const float2 e0 = add(d0, d2), e2 = sub(d0, d2), e1 = add(d1, d3), e3 = mi(sub(d1, d3));
            const float2 y[8] = {add(c0, c1), add(e0, e1), add(c2, c3), add(e2, e3),
                                 sub(c0, c1), sub(e0, e1), sub(c2, c3), sub(e2, e3)};
            const int d = ((j - k) << 3) + k;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (SNK && last)
                    snk(d + q * Ns, y[q]);
                else
                    buf[a ^ 1][d + q * Ns] = y[q];
            }
        }
        __syncthreads();
        a ^= 1;
    }
    if (Ns * 4 == N) {
        constexpr int Q = N / 4;
        const int step = N / (Ns * 4);
        for (int j = threadIdx.x; j < Q; j += 256) {
            const int k = j & (Ns - 1);
            const float2 v0 = buf[a][j];
            const float2 v1 = cmul(buf[a][j + Q], twiddle<N>(tw, k * step));
            const float2 v2 = cmul(buf[a][j + 2 * Q], twiddle<N>(tw, 2 * k * step));
            const float2 v3 = cmul(buf[a][j + 3 * Q], twiddle<N>(tw, 3 * k * step));
            const int d = ((j - k) << 2) + k;
            const float2 y[4] = {make_float2(v0.x + v1.x + v2.x + v3.x, v0.y + v1.y + v2.y + v3.y),
                                 make_float2(v0.x + v1.y - v2.x - v3.y, v0.y - v1.x - v2.y + v3.x),
                                 make_float2(v0.x - v1.x + v2.x - v3.x, v0.y - v1.y + v2.y - v3.y),
                                 make_float2(v0.x - v1.y - v2.x + v3.y, v0.y + v1.x - v2.y - v3.x)};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (SNK)
                    snk(d + q * Ns, y[q]);
                else
                    buf[a ^ 1][d + q * Ns] = y[q];
            }
        }
        __syncthreads();
        a ^= 1;
    } else if (Ns * 2 == N) {
        constexpr int H2 = N / 2;
        const int step = N / (Ns * 2);
        for (int j = threadIdx.x; j < H2; j += 256) {
            const int k = j & (Ns - 1);
            const float2 v0 = buf[a][j], v1 = cmul(buf[a][j + H2], twiddle<N>(tw, k * step));
            const int d = ((j - k) << 1) + k;
            const float2 y0 = make_float2(v0.x + v1.x, v0.y + v1.y), y1 = make_float2(v0.x - v1.x, v0.y - v1.y);
            if (SNK) {
                snk(d, y0);
                snk(d + Ns, y1);
            } else {
                buf[a ^ 1][d] = y0;
                buf[a ^ 1][d + Ns] = y1;
            }
        }
        __syncthreads();
        a ^= 1;
    }
    return a;
}

template <int N>
__device__ __forceinline__ int fft_stockham(float2 (*buf)[N], int a, const float2* __restrict__ tw)
{
    auto none_src = [](int, int, int) { return make_float2(0.0f, 0.0f); };
    auto none_snk = [](int, float2) {};
    return fft_stockham<N, false, false>(buf, a, tw, none_src, none_snk);
}

template <int LOGN>
__global__ __launch_bounds__(256) void k_fdk_filter(const float* __restrict__ in, float* __restrict__ out,
                                                    const float* __restrict__ h, long long rows, int na)
{
    constexpr int N = 1 << LOGN;
    constexpr int RPT = (N + 255) / 256;   // response values per thread
    __shared__ float2 buf[2][N];
    __shared__ float2 tw[N / 2];
    const int tid = threadIdx.x;
    for (int t = tid; t < N / 2; t += 256) {
        float sn, cs;
        sincospif(-2.0f * (float)t / (float)N, &sn, &cs);
        tw[t] = make_float2(cs, sn);
    }
    // the taps laid out circularly: offset d at index d mod N
    for (int i = tid; i < N; i += 256) {
        float v = 0.0f;
        if (i <= na - 1)
            v = h[na - 1 + i];
        else if (i >= N - (na - 1))
            v = h[na - 1 + i - N];
        buf[0][i] = make_float2(v, 0.0f);
    }
    __syncthreads();
    const int sr = fft_stockham<N>(buf, 0, tw);
    // the response stays in registers: thread tid multiplies entries
    // tid + 256 j, and LDS holds only the two work buffers and half the
    // twiddles (4 workgroups per CU at N = 2048 instead of 2)
    float2 resp[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int i = tid + 256 * j;
        resp[j] = i < N ? buf[sr][i] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const float inv_n = 1.0f / (float)N;
    for (long long pr = blockIdx.x; 2 * pr < rows; pr += gridDim.x) {
        const long long r1 = 2 * pr, r2 = r1 + 1;
        const bool has2 = r2 < rows;
        if constexpr (N >= 2048) {
            // 8 LDS passes per row pair: the forward transform's first pass
            // reads the rows from global memory (indices >= N/2 > na - 1 are
            // the zero padding), the inverse's first pass multiplies by the
            // response held in registers (element tid + 256 (u + m N/2048)),
            // and its last pass writes the filtered rows out
            constexpr int E = N / 8;
            auto load = [&](int, int m, int j) {
                const int i = j + m * E;
                float x1 = 0.0f, x2 = 0.0f;
                if (i < na) {
                    x1 = in[r1 * na + i];
                    if (has2) x2 = in[r2 * na + i];
                }
                return make_float2(x1, x2);
            };
            auto none_snk = [](int, float2) {};
            const int s1 = fft_stockham<N, true, false>(buf, 0, tw, load, none_snk);
            auto mulh = [&](int u, int m, int j) {   // conj(Z H)
                const float2 z = buf[s1][j + m * E], hh = resp[u + m * (E / 256)];
                return make_float2(z.x * hh.x - z.y * hh.y, -(z.x * hh.y + z.y * hh.x));
            };
            auto store = [&](int i, float2 yv) {
                if (i < na) {
                    out[r1 * na + i] = yv.x * inv_n;
                    if (has2) out[r2 * na + i] = -yv.y * inv_n;
                }
            };
            fft_stockham<N, true, true>(buf, s1, tw, mulh, store);
        } else {
            for (int i = tid; i < N; i += 256) {
                float x1 = 0.0f, x2 = 0.0f;
                if (i < na) {
                    x1 = in[r1 * na + i];
                    if (has2) x2 = in[r2 * na + i];
                }
                buf[0][i] = make_float2(x1, x2);
            }
            __syncthreads();
            const int s1 = fft_stockham<N>(buf, 0, tw);
#pragma unroll
            for (int j = 0; j < RPT; ++j) {   // conj(Z H)
                const int i = tid + 256 * j;
                if (i < N) {
                    const float2 z = buf[s1][i], hh = resp[j];
                    buf[s1][i] = make_float2(z.x * hh.x - z.y * hh.y, -(z.x * hh.y + z.y * hh.x));
                }
            }
            __syncthreads();
            const int s2 = fft_stockham<N>(buf, s1, tw);
            for (int i = tid; i < na; i += 256) {
                const float2 y = buf[s2][i];
                out[r1 * na + i] = y.x * inv_n;
                if (has2) out[r2 * na + i] = -y.y * inv_n;
            }
        }
        __syncthreads();   // the next pair overwrites buf[0]
    }
}

// Voxel-driven weighted backprojection.  Thread = (ix, iy, z-chunk of ZC
// voxels); all views accumulate in registers; one write per voxel.  Per
// (x, y, view) the detector column, magnification and distance weight are
// computed once in FP64 (reciprocals instead of divisions) and shared by the
// ZC voxels of the column; the row coordinate is FP32 per voxel, from the
// chunk's first voxel and the per-voxel step: fb_k = fma(k, vz mb, fb_first),
// one FMA per voxel and no per-voxel z in registers.  With
// PAIR (na >= 2, so ia0 <= na - 2) the two taps of a detector row are read as
// one 8-byte load (4-byte aligned; gfx950 loads need only dword alignment).
constexpr int ZC = 16;

template <bool PAIR>
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
    const float zc = (float)(z0 + (izc + 0.5) * vz), vzf = (float)vz;   // first voxel's z, step
    const int kn = min(ZC, nz - izc);
    const double ha = na * da * 0.5, hb = nb * db * 0.5;
    const double inv_da = 1.0 / da;
    const float inv_db = (float)(1.0 / db);
    const float fb0 = (float)(hb / db - bi);
    const float nb1 = (float)(nb - 1);
    const unsigned rowstep = (nb >= 2) ? 4u * (unsigned)na : 0u;   // byte offset of row ib0 + 1
    for (int iv = 0; iv < nv; ++iv) {
        // view-uniform values first, before the lanes diverge
        const float* qv = q + (size_t)nb * na * iv;
        const float off = offsets ? offsets[iv] : 0.0f;
        const double c = trig[2 * iv], s = trig[2 * iv + 1];
        const double U = SID + x * c + y * s;
        if (!(U > 0.0)) continue;
        const double iU = rcp64(U);
        const double mag = SAD * iU;
        const double fa = ((-x * s + y * c) * mag + ha) * inv_da - ai;
        if (!(fa >= 0.0 && fa <= (double)(na - 1))) continue;
        const int ia0 = min((int)floor(fa), max(na - 2, 0));
        const float ta = (float)(fa - ia0);
        const float wr = (float)(SID * iU);
        const float w = wr * wr;
        const float mb = (float)mag * inv_db;
        const int ia1 = min(ia0 + 1, na - 1);
        // branch-free over the ZC voxels, loads first, so all of them are in
        // flight together: a row coordinate outside the detector is clamped for
        // the address and its sample dropped by the select
        float2 p0[ZC], p1[ZC];
        float tb[ZC];
        bool valid[ZC];
        const float fbv = fmaf(-off, mb, fb0);   // row coordinate of z = 0
        const float fbs = fmaf(zc, mb, fbv), dzb = no_contract(vzf * mb);
        // a full chunk whose row coordinates all lie in [0, nb - 1) (ends
        // bound a monotone fma): no clamp, test or select, the same sums
        const float fbl = fmaf((float)(ZC - 1), dzb, fbs);
        if (PAIR && kn == ZC && fminf(fbs, fbl) >= 0.0f && fmaxf(fbs, fbl) < nb1) {
            const char* qb = reinterpret_cast<const char*>(qv);
#pragma unroll
            for (int k = 0; k < ZC; ++k) {
                const float fb = fmaf((float)k, dzb, fbs);
                const int ib0 = (int)fb;
                tb[k] = fb - (float)ib0;
                const unsigned o0 = 4u * (unsigned)(__mul24(ib0, na) + ia0), o1 = o0 + rowstep;
                __builtin_memcpy(&p0[k], qb + o0, 8);
                __builtin_memcpy(&p1[k], qb + o1, 8);
            }
#pragma unroll
            for (int k = 0; k < ZC; ++k) {
                const float top = fmaf(ta, p0[k].y - p0[k].x, p0[k].x);
                const float bot = fmaf(ta, p1[k].y - p1[k].x, p1[k].x);
                acc[k] = fmaf(w, fmaf(tb[k], bot - top, top), acc[k]);
            }
            continue;
        }
#pragma unroll
        for (int k = 0; k < ZC; ++k) {
            const float fb = fmaf((float)k, dzb, fbs);
            valid[k] = k < kn && fb >= 0.0f && fb <= nb1;
            const float fbc = __builtin_amdgcn_fmed3f(fb, 0.0f, nb1);
            const int ib0 = min((int)fbc, max(nb - 2, 0));   // fbc >= 0: truncation is floor
            const int ib1 = min(ib0 + 1, nb - 1);
            tb[k] = fbc - (float)ib0;
            if (PAIR) {
                // 32-bit byte offsets from the view's base (a view has < 2^29
                // cells; 24-bit multiplies, na and nb < 2^24)
                const unsigned o0 = 4u * (unsigned)(__mul24(ib0, na) + ia0), o1 = o0 + rowstep;
                const char* qb = reinterpret_cast<const char*>(qv);
                __builtin_memcpy(&p0[k], qb + o0, 8);
                __builtin_memcpy(&p1[k], qb + o1, 8);
            } else {
                const float* r0p = qv + (size_t)ib0 * na;
                const float* r1p = qv + (size_t)ib1 * na;
                p0[k] = make_float2(r0p[ia0], r0p[ia1]);
                p1[k] = make_float2(r1p[ia0], r1p[ia1]);
            }
        }
#pragma unroll
        for (int k = 0; k < ZC; ++k) {
            const float top = fmaf(ta, p0[k].y - p0[k].x, p0[k].x);
            const float bot = fmaf(ta, p1[k].y - p1[k].x, p1[k].x);
            const float sum = fmaf(w, fmaf(tb[k], bot - top, top), acc[k]);
            acc[k] = valid[k] ? sum : acc[k];
        }
    }
#pragma unroll
    for (int k = 0; k < ZC; ++k) {
        const int iz = izc + k;
        if (iz < nz) image[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)] = acc[k];
    }
}

// Staged variants of k_fdk_bp.  A workgroup owns BXS x BYS voxel columns
// times SZ voxels in z (a thread: CPT columns' SZ voxels each).  Per view it
// copies the block's detector footprint into LDS once (the bounding rectangle
// of the projected voxel-centre box with one cell of margin) and reads every
// voxel's four taps from there, instead of gathering 16 bytes per voxel and
// view through the vector cache.  The footprint of view iv + 1 is loaded into
// registers before view iv is computed and written to the other LDS buffer
// after it: one barrier per view.  Per voxel and view the arithmetic is
// k_fdk_bp's, operation for operation, so all mappings give identical images.
// A view whose footprint exceeds RW x RH cells is gathered from global
// memory, as k_fdk_bp does.  Footprints are bounded only when every voxel
// centre lies in front of the source (U > 0 for every view), which the
// launcher checks.
//
// Each detector cell of a view is read by every block its ray crosses, so the
// bytes a view asks of the L2 fall with the block's side: 16 x 16 columns
// (4 waves, 4 workgroups per CU: the default) or 64 x 64 columns (8 waves,
// eight columns per thread, one workgroup and 130 KB of LDS per CU, blocks
// dealt to the XCDs in contiguous ranges: a quarter of the footprint bytes
// and a third of the HBM traffic, but slower, since one workgroup per CU
// leaves nothing to run while its waves wait at the per-view barrier).
//
// LDS rows have a fixed stride of RW + 4 floats: a voxel's four taps are two
// ds_read2_b32 at constant offsets (0, 1) and (SRS, SRS + 1) from one
// address.  Staging: float4 slot j of thread t is float4 t + j T of the
// footprint (T threads, RW / 4 float4 per row), one 16-byte load whose
// address is a per-view scalar plus a per-thread constant; the lanes of a
// wave cover whole rows, or runs of 16 float4 of 4 rows.
constexpr int SZ = 16;          // voxels per thread along z

struct Foot {
    int ua0, vb0, wq, H;        // first column and row, row length / 4, rows; wq == 0: not staged
};

// Footprints of views v0 .. v0 + 7 of the block whose voxel-centre box is
// [cx0, cx1] x [cy0, cy1] x [cz0, cz1]: lane group g = lane / 8 takes view
// v0 + g (views past nv give an unstaged footprint), lane & 7 a corner.  The
// projective map sends the box into the hull of its corners' images, so the
// corners' extreme indices bound every voxel's taps; the one-cell margin
// covers FP32 rounding of the voxels' row coordinates (~1e-3 cells).  A
// footprint wider than rw or taller than rh cells is not staged.
// lane l's value of a double held per lane
__device__ __forceinline__ double read_lane64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <bool PIX, bool CYL = false>
__device__ int4 fdk_footprints(const double* __restrict__ trig, const float* __restrict__ angles,
                               const float* __restrict__ offsets, int nv, int v0, double cx0, double cx1,
                               double cy0, double cy1, double cz0, double cz1, double SID, double SAD, double ha,
                               double hb, double inv_da, double ai, double inv_db, double bi, int na, int nb, int rw,
                               int rh)
{
    const int lane = threadIdx.x & 63, q = lane & 7;
    const int iv = min(v0 + (lane >> 3), nv - 1);
    double c, s;
    if (PIX) {
        const double a = (double)angles[iv];
        c = cos(a);
        s = sin(a);
    } else {
        c = trig[2 * iv];
        s = trig[2 * iv + 1];
    }
    const double off = offsets ? (double)offsets[iv] : 0.0;
    const double x = (q & 1) ? cx1 : cx0, y = (q & 2) ? cy1 : cy0, z = (q & 4) ? cz1 : cz0;
    const double U = SID + x * c + y * s;
    const double mag = SAD * rcp64(U);
    // the launcher guarantees U >= 1e-3 SID, so both indices are finite
    // CYL: the column is the ray's in-plane angle, monotone in the slope
    // y'/U, whose extremes over the box lie on its corners as well
    float amn = CYL ? (float)((double)atan2f((float)(-x * s + y * c), (float)U) * inv_da + ha * inv_da - ai)
                    : (float)(((-x * s + y * c) * mag + ha) * inv_da - ai);
    float bmn = (float)(((z - off) * mag + hb) * inv_db - bi);
    float amx = amn, bmx = bmn;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        amn = fminf(amn, __shfl_xor(amn, m));
        amx = fmaxf(amx, __shfl_xor(amx, m));
        bmn = fminf(bmn, __shfl_xor(bmn, m));
        bmx = fmaxf(bmx, __shfl_xor(bmx, m));
    }
    // index bounds clamped in float before the conversion
    int ua0 = (int)fminf(fmaxf(floorf(amn) - 1.0f, 0.0f), (float)(na - 2));
    int ua1 = (int)fminf(fmaxf(floorf(amx) + 2.0f, 0.0f), (float)(na - 1));
    ua1 = max(ua1, ua0 + 1);
    const int vb0 = (int)fminf(fmaxf(floorf(bmn) - 1.0f, 0.0f), (float)(nb - 2));
    int vb1 = (int)fminf(fmaxf(floorf(bmx) + 2.0f, 0.0f), (float)(nb - 1));
    vb1 = max(vb1, vb0 + 1);
    // whole 16-byte groups: the row widened to a multiple of 4 cells, moved
    // left where it would pass the detector's last column
    const int ws = (ua1 - ua0 + 4) & ~3;
    ua0 = min(ua0, na - ws);
    const int H = vb1 - vb0 + 1;
    const bool staged = v0 + (lane >> 3) < nv && ua0 >= 0 && ws <= rw && H <= rh;
    return make_int4(ua0, vb0, staged ? ws >> 2 : 0, H);
}

// 16 bytes from a 4-byte aligned address (gfx950 global loads need only dword
// alignment), returned by value so the destination stays in registers
__device__ __forceinline__ float4 load16(const char* p)
{
    float4 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}

__device__ __forceinline__ Foot foot_of(const int4& grp, int g)
{
    Foot f;
    f.ua0 = __builtin_amdgcn_readlane(grp.x, 8 * g);
    f.vb0 = __builtin_amdgcn_readlane(grp.y, 8 * g);
    f.wq = __builtin_amdgcn_readlane(grp.z, 8 * g);
    f.H = __builtin_amdgcn_readlane(grp.w, 8 * g);
    return f;
}

template <int BXS, int BYS, int NWV, int RW, int RH>
struct StagedShape {
    static constexpr int T = 64 * NWV;                 // threads
    static constexpr int CPT = BXS * BYS / T;          // voxel columns per thread
    static constexpr int SRS = RW + 4;                 // LDS row stride (floats)
    static constexpr int R4 = RW / 4;                  // float4 per footprint row
    static constexpr int SNJ = RH * R4 / T;            // staging slots per thread and view
    static_assert(CPT * T == BXS * BYS && SNJ * T == RH * R4 && RW % 4 == 0, "block shape");
    static_assert(T % R4 == 0, "staging columns the same for every slot");
};

template <int BXS, int BYS, int NWV, int RW, int RH, bool XCDMAP, bool PIX = false, bool CYL = false>
__global__ __launch_bounds__(64 * NWV, (RW > 64 ? 1 : 16 / NWV)) void k_fdk_bp_staged(
    const float* __restrict__ q, const double* __restrict__ trig, const float* __restrict__ offsets, int nv, int nx,
    int ny, int nz, double x0, double y0, double z0, double vx, double vy, double vz, double SID, double SAD, int na,
    double da, double ai, int nb, double db, double bi, float* __restrict__ image,
    const float* __restrict__ angles = nullptr)
{
    using S = StagedShape<BXS, BYS, NWV, RW, RH>;
    constexpr int CPT = S::CPT, SRS = S::SRS, SNJ = S::SNJ, T = S::T, R4 = S::R4;
    __shared__ float4 lds4[2][RH * SRS / 4];
    const int tid = threadIdx.x;
    // Block coordinates.  XCDMAP: a 1-D grid (a multiple of 8) whose
    // workgroups the hardware deals to the 8 XCDs in turn (id mod 8) is
    // remapped so that each XCD gets a contiguous eighth of the blocks (x
    // fastest, then y, then z): the blocks an XCD runs at once are
    // neighbours in one z slab, whose footprints share detector rows in
    // that XCD's L2.
    int bxi, byi, bzi;
    if (XCDMAP) {
        const int nbx = (nx + BXS - 1) / BXS, nby = (ny + BYS - 1) / BYS;
        const int total = nbx * nby * ((nz + SZ - 1) / SZ);
        const int per = (int)(gridDim.x >> 3);
        const int n = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
        if (n >= total) return;                         // the whole workgroup: no barrier reached
        bxi = n % nbx;
        byi = (n / nbx) % nby;
        bzi = n / (nbx * nby);
    } else {
        bxi = blockIdx.x;
        byi = blockIdx.y;
        bzi = blockIdx.z;
    }
    const int izc = bzi * SZ;
    const int kn = min(SZ, nz - izc);
    // this thread's columns: column index tid + c T of the block, x fastest;
    // T is a multiple of BXS, so all of them share x and step y by T / BXS
    static_assert(T % BXS == 0, "columns of a thread share x");
    constexpr int YSTEP = T / BXS;
    const int ix = bxi * BXS + tid % BXS;
    const int iy0 = byi * BYS + tid / BXS;
    const double x = x0 + (ix + 0.5) * vx;
    float acc[CPT][SZ];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
#pragma unroll
        for (int k = 0; k < SZ; ++k) acc[c][k] = 0.0f;
    }
    const float zc = (float)(z0 + (izc + 0.5) * vz), vzf = (float)vz;   // first voxel's z, step
    // PIX: the voxel-driven op's detector indices, y' mag / da + na/2 (integer
    // na/2) - ai (the launcher passes ai + 0.5, bi + 0.5)
    const double ha = PIX ? (double)(na / 2) * da : na * da * 0.5;
    const double hb = PIX ? (double)(nb / 2) * db : nb * db * 0.5;
    const double inv_da = 1.0 / da, inv_db64 = 1.0 / db;
    const float inv_db = (float)inv_db64;
    const float fb0 = (float)(hb / db - bi);
    const float nb1 = (float)(nb - 1);
    const unsigned rowstep = 4u * (unsigned)na;
    // PIX: cos / sin of views v0 + lane of the current 64-view group, per lane
    double cs_l = 0.0, sn_l = 0.0;
    auto trig_of = [&](int iv, double& c, double& sn) {
        if (PIX) {
            if ((iv & 63) == 0) {
                const double a = (double)angles[min(iv + (tid & 63), nv - 1)];
                cs_l = cos(a);
                sn_l = sin(a);
            }
            c = read_lane64(cs_l, iv & 63);
            sn = read_lane64(sn_l, iv & 63);
        } else {
            c = trig[2 * iv];
            sn = trig[2 * iv + 1];
        }
    };
    // the block's voxel-centre box
    const int bx1 = min(bxi * BXS + BXS - 1, nx - 1), by1 = min(byi * BYS + BYS - 1, ny - 1);
    const double cx0 = x0 + (bxi * BXS + 0.5) * vx, cx1 = x0 + (bx1 + 0.5) * vx;
    const double cy0 = y0 + (byi * BYS + 0.5) * vy, cy1 = y0 + (by1 + 0.5) * vy;
    // (the box spans all SZ z slots, also those past the volume, whose samples
    // are dropped but whose clamped taps are read)
    const double cz0 = z0 + (izc + 0.5) * vz, cz1 = z0 + (izc + SZ - 1 + 0.5) * vz;
    auto footprints = [&](int v0) {
        return fdk_footprints<PIX, CYL>(trig, angles, offsets, nv, v0, cx0, cx1, cy0, cy1, cz0, cz1, SID, SAD, ha, hb,
                                   inv_da, ai, inv_db64, bi, na, nb, RW, RH);
    };
    // this thread's staging cells: float4 tid + j T of the footprint, row
    // (tid + j T) / R4, cells 4 m .. 4 m + 3 with m = tid % R4 (T is a
    // multiple of R4, so m is the same for every j)
    const int m4 = 4 * (tid % R4);
    const int rsub = tid / R4;                          // row of slot 0
    constexpr int RSTEP = T / R4;                       // rows between slots (R4 <= T)
    static_assert(R4 <= T, "one footprint row per slot set");
    const unsigned goff = 4u * (unsigned)(rsub * na + m4);    // bytes, within the view (< 2^31)
    const int soff = rsub * SRS + m4;                         // floats, within an LDS buffer
    auto stage_load = [&](const Foot& f, int jv, float4 (&st)[SNJ]) {
        if (f.wq == 0) return;
        const bool col_ok = m4 < 4 * f.wq;
        const char* qb = reinterpret_cast<const char*>(q + (size_t)nb * na * jv);
#pragma unroll
        for (int j = 0; j < SNJ; ++j) {
            const int r0 = RSTEP * j;
            if (col_ok && r0 + rsub < f.H) st[j] = load16(qb + 4u * (unsigned)((f.vb0 + r0) * na + f.ua0) + goff);
        }
    };
    auto stage_store = [&](const Foot& f, int b, const float4 (&st)[SNJ]) {
        if (f.wq == 0) return;
        const bool col_ok = m4 < 4 * f.wq;
        float* lb = reinterpret_cast<float*>(&lds4[b][0]);
#pragma unroll
        for (int j = 0; j < SNJ; ++j) {
            const int r0 = RSTEP * j;
            if (col_ok && r0 + rsub < f.H) *reinterpret_cast<float4*>(lb + r0 * SRS + soff) = st[j];
        }
    };
    int4 grp = footprints(0);
    Foot cur = foot_of(grp, 0);
    {
        float4 st[SNJ];
        stage_load(cur, 0, st);
        stage_store(cur, 0, st);
    }
    __syncthreads();
    for (int iv = 0; iv < nv; ++iv) {
        const int jn = iv + 1;
        Foot nxt = {0, 0, 0, 0};
        if (jn < nv) {
            if ((jn & 7) == 0) grp = footprints(jn);
            nxt = foot_of(grp, jn & 7);
        }
        float4 st[SNJ];
        stage_load(nxt, jn, st);
        // view iv, from LDS buffer iv & 1 (or from global memory)
        const float* qv = q + (size_t)nb * na * iv;
        const float off = offsets ? offsets[iv] : 0.0f;
        double cth, sth;
        trig_of(iv, cth, sth);
        // The next view's cells go to the other buffer (nobody reads it before
        // the barrier below) once the first column is done: with several
        // columns per thread their registers are then free for the others.
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            do {   // one column's SZ samples; break leaves it
                const int iy = iy0 + c * YSTEP;
                const double y = y0 + (iy + 0.5) * vy;
                const double U = SID + x * cth + y * sth;
                if (!(ix < nx && iy < ny && U > 0.0)) break;
                const double iU = rcp64(U);
                const double mag = SAD * iU;
                // CYL (cylindrical detector, the voxel-driven op only): the
                // column from the ray's in-plane angle, atan2 in FP32 as the
                // op computes it (pixel.cu.cc:70-71)
                const double fa = CYL ? (double)atan2f((float)(-x * sth + y * cth), (float)U) * inv_da + ha * inv_da - ai
                                      : ((-x * sth + y * cth) * mag + ha) * inv_da - ai;
                // PIX: 0 <= a < na - 1 (the op skips a view whose 2 x 2
                // neighbourhood leaves the detector, pixel.cu.cc:43-50)
                if (!(fa >= 0.0 && (PIX ? fa < (double)(na - 1) : fa <= (double)(na - 1)))) break;
                const int ia0 = min((int)floor(fa), na - 2);
                const float ta = (float)(fa - ia0);
                const float wr = (float)(SID * iU);
                const float w = PIX ? 1.0f : wr * wr;
                const float mb = (float)mag * inv_db;
                const float fbv = fmaf(-off, mb, fb0);   // row coordinate of z = 0
                const float fbs = fmaf(zc, mb, fbv), dzb = no_contract(vzf * mb);
                if (cur.wq > 0) {
                    // float index of row ib0's tap ia0: ib0 SRS + base
                    const float* lb = reinterpret_cast<const float*>(&lds4[iv & 1][0]);
                    const int base = (ia0 - cur.ua0) - cur.vb0 * SRS;
                    // A column whose SZ row coordinates all lie in [0, nb - 1)
                    // (fma is monotone in k: its ends bound it) needs no clamp,
                    // no validity test and no select: the same sums, bitwise.
                    const float fbl = fmaf((float)(SZ - 1), dzb, fbs);
                    if (kn == SZ && fminf(fbs, fbl) >= 0.0f && fmaxf(fbs, fbl) < nb1) {
#pragma unroll
                        for (int k = 0; k < SZ; ++k) {
                            const float fb = fmaf((float)k, dzb, fbs);
                            const int ib0 = (int)fb;                  // 0 <= fb < nb - 1
                            const float tb = fb - (float)ib0;
                            const float* p = lb + (__mul24(ib0, SRS) + base);
                            const float p00 = p[0], p01 = p[1], p10 = p[SRS], p11 = p[SRS + 1];
                            const float top = fmaf(ta, p01 - p00, p00);
                            const float bot = fmaf(ta, p11 - p10, p10);
                            acc[c][k] = fmaf(w, fmaf(tb, bot - top, top), acc[c][k]);
                        }
                        break;
                    }
#pragma unroll
                    for (int k = 0; k < SZ; ++k) {
                        const float fb = fmaf((float)k, dzb, fbs);
                        const float fbc = __builtin_amdgcn_fmed3f(fb, 0.0f, nb1);
                        // 0 <= fb <= nb - 1 (PIX: < nb - 1)
                        const bool valid = k < kn && fb == fbc && (!PIX || fb < nb1);
                        const int ib0 = min((int)fbc, nb - 2);    // fbc >= 0: truncation is floor
                        const float tb = fbc - (float)ib0;
                        const float* p = lb + (__mul24(ib0, SRS) + base);
                        const float p00 = p[0], p01 = p[1], p10 = p[SRS], p11 = p[SRS + 1];
                        const float top = fmaf(ta, p01 - p00, p00);
                        const float bot = fmaf(ta, p11 - p10, p10);
                        const float sum = fmaf(w, fmaf(tb, bot - top, top), acc[c][k]);
                        acc[c][k] = valid ? sum : acc[c][k];
                    }
                } else {
                    const char* qb = reinterpret_cast<const char*>(qv);
#pragma unroll
                    for (int k = 0; k < SZ; ++k) {
                        const float fb = fmaf((float)k, dzb, fbs);
                        const float fbc = __builtin_amdgcn_fmed3f(fb, 0.0f, nb1);
                        const bool valid = k < kn && fb == fbc && (!PIX || fb < nb1);
                        const int ib0 = min((int)fbc, nb - 2);
                        const float tb = fbc - (float)ib0;
                        const unsigned o0 = 4u * (unsigned)(__mul24(ib0, na) + ia0), o1 = o0 + rowstep;
                        float2 p0, p1;
                        __builtin_memcpy(&p0, qb + o0, 8);
                        __builtin_memcpy(&p1, qb + o1, 8);
                        const float top = fmaf(ta, p0.y - p0.x, p0.x);
                        const float bot = fmaf(ta, p1.y - p1.x, p1.x);
                        const float sum = fmaf(w, fmaf(tb, bot - top, top), acc[c][k]);
                        acc[c][k] = valid ? sum : acc[c][k];
                        // several columns' accumulators leave no registers for
                        // every sample's taps in flight at once
                        if (CPT > 1 && k % 4 == 3) __builtin_amdgcn_sched_barrier(0);
                    }
                }
            } while (0);
            if (c == 0) stage_store(nxt, jn & 1, st);
        }
        __syncthreads();
        cur = nxt;
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int iy = iy0 + c * YSTEP;
        if (!(ix < nx && iy < ny)) continue;
#pragma unroll
        for (int k = 0; k < SZ; ++k) {
            if (k < kn) image[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * (size_t)(izc + k))] = acc[c][k];
        }
    }
}

// the two staged shapes: 16 x 16 columns with footprints up to 64 x 64 cells,
// 64 x 64 columns with footprints up to 256 x 64 cells
constexpr int WBX = 64, WBY = 64;   // the wide blocks' columns
#define NEFCT_FDK_SMALL 16, 16, 4, 64, 64, false
#define NEFCT_FDK_WIDE WBX, WBY, 8, 256, 64, true

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

namespace nefct_internal {

bool pixel_bp_staged(bool cyl, const float* pv, const int grid[3], const float center[3], const float size[3],
                     const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb, float db,
                     float bi, float* image, hipStream_t stream)
{
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    if (na < 2 || nb < 2 || na >= (1 << 23) || nb >= (1 << 23) || (size_t)na * (size_t)nb >= (1ull << 29))
        return false;
    if ((ny + 15) / 16 > 65535 || (nz + SZ - 1) / SZ > 65535) return false;
    // the op's voxel centres: (i + 0.5) size / n + center
    const double vx = (double)size[0] / nx, vy = (double)size[1] / ny, vz = (double)size[2] / nz;
    const double x0 = center[0], y0 = center[1], z0 = center[2];
    double r2 = 0.0;
    for (int cxy = 0; cxy < 4; ++cxy) {
        const double px = x0 + ((cxy & 1) ? nx - 0.5 : 0.5) * vx, py = y0 + ((cxy & 2) ? ny - 0.5 : 0.5) * vy;
        r2 = std::max(r2, px * px + py * py);
    }
    if (!(std::sqrt(r2) <= 0.999 * (double)SID)) return false;
    const dim3 g((nx + 15) / 16, (ny + 15) / 16, (nz + SZ - 1) / SZ);
    if (cyl)
        hipLaunchKernelGGL((k_fdk_bp_staged<NEFCT_FDK_SMALL, true, true>), g, dim3(256), 0, stream, pv, nullptr,
                           nullptr, nv, nx, ny, nz, x0, y0, z0, vx, vy, vz, (double)SID, (double)SAD, na, (double)da,
                           (double)ai + 0.5, nb, (double)db, (double)bi + 0.5, image, angles);
    else
        hipLaunchKernelGGL((k_fdk_bp_staged<NEFCT_FDK_SMALL, true>), g, dim3(256), 0, stream, pv, nullptr, nullptr,
                           nv, nx, ny, nz, x0, y0, z0, vx, vy, vz, (double)SID, (double)SAD, na, (double)da,
                           (double)ai + 0.5, nb, (double)db, (double)bi + 0.5, image, angles);
    return true;
}

}  // namespace nefct_internal

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
    hipLaunchKernelGGL(k_fdk_preweight, dim3((na + 255) / 256, nv, (nb + PW_ROWS - 1) / PW_ROWS), dim3(256), 0,
                       (hipStream_t)stream,
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
    int logn = 6;
    while ((1 << logn) < 2 * na - 1) ++logn;
    const long long pairs = ((long long)rows + 1) / 2;
    const dim3 g((unsigned)(pairs < 4096 ? pairs : 4096));
    hipStream_t s = (hipStream_t)stream;
    switch (logn) {
    case 6: hipLaunchKernelGGL(k_fdk_filter<6>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    case 7: hipLaunchKernelGGL(k_fdk_filter<7>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    case 8: hipLaunchKernelGGL(k_fdk_filter<8>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    case 9: hipLaunchKernelGGL(k_fdk_filter<9>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    case 10: hipLaunchKernelGGL(k_fdk_filter<10>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    case 11: hipLaunchKernelGGL(k_fdk_filter<11>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    default: hipLaunchKernelGGL(k_fdk_filter<12>, g, dim3(256), 0, s, in, out, taps, (long long)rows, na); break;
    }
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

size_t nefct_fdk_workspace_size(int nv) { return sizeof(double) * 2 * (size_t)(nv > 0 ? nv : 1); }

int nefct_fdk_backproject_mapped(const float* filtered, const int grid[3], const float center[3],
                                 const float size[3], const float* angles, const float* offsets, int nv, float SID,
                                 float SAD, int na, float da, float ai, int nb, float db, float bi, float* image,
                                 int mapping, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!filtered || !grid || !center || !size || !angles || !image) return NEFCT_ERR_NULL;
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0 || nv <= 0 || na <= 0 || nb <= 0) return NEFCT_ERR_SHAPE;
    if ((grid[1] + 3) / 4 > 65535 || (grid[2] + ZC - 1) / ZC > 65535) return NEFCT_ERR_SHAPE;
    // 32-bit byte offsets within a view, 24-bit row multiplies
    // na, nb < 2^23: __mul24 takes signed 24-bit operands (row offsets ib na < 2^29)
    if ((size_t)na * (size_t)nb >= (1ull << 29) || na >= (1 << 23) || nb >= (1 << 23)) return NEFCT_ERR_SHAPE;
    if (!finite_pos(SID) || !finite_pos(SAD) || !finite_pos(da) || !finite_pos(db) || !finite_pos(size[0]) ||
        !finite_pos(size[1]) || !finite_pos(size[2]) || !std::isfinite(center[0]) || !std::isfinite(center[1]) ||
        !std::isfinite(center[2]) || !std::isfinite(ai) || !std::isfinite(bi))
        return NEFCT_ERR_GEOMETRY;
    if (mapping != NEFCT_FDK_BP_AUTO && mapping != NEFCT_FDK_BP_GATHER && mapping != NEFCT_FDK_BP_STAGED &&
        mapping != NEFCT_FDK_BP_WIDE)
        return NEFCT_ERR_UNSUPPORTED;
    if (!workspace || workspace_bytes < nefct_fdk_workspace_size(nv)) return NEFCT_ERR_WORKSPACE;
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    const double vx = (double)size[0] / nx, vy = (double)size[1] / ny, vz = (double)size[2] / nz;
    const double x0 = (double)center[0] - 0.5 * size[0], y0 = (double)center[1] - 0.5 * size[1],
                 z0 = (double)center[2] - 0.5 * size[2];
    // The staged kernels need bounded footprints: every voxel centre at least
    // 1e-3 SID in front of the source at every angle (U >= SID - r >= 1e-3 SID
    // with r the largest in-plane distance of a voxel centre from the axis).
    double r2 = 0.0;
    for (int cxy = 0; cxy < 4; ++cxy) {
        const double px = x0 + ((cxy & 1) ? nx - 0.5 : 0.5) * vx, py = y0 + ((cxy & 2) ? ny - 0.5 : 0.5) * vy;
        r2 = std::max(r2, px * px + py * py);
    }
    const bool stageable = na >= 2 && nb >= 2 && std::sqrt(r2) <= 0.999 * (double)SID &&
                           (ny + 15) / 16 <= 65535 && (nz + SZ - 1) / SZ <= 65535;
    // the wide mapping's 1-D grid: block counts in int
    const long long wide_blocks = (long long)((nx + WBX - 1) / WBX) * ((ny + WBY - 1) / WBY) * ((nz + SZ - 1) / SZ);
    const bool wide_ok = wide_blocks <= (1ll << 30);
    if ((mapping == NEFCT_FDK_BP_STAGED || mapping == NEFCT_FDK_BP_WIDE) && !stageable) return NEFCT_ERR_UNSUPPORTED;
    if (mapping == NEFCT_FDK_BP_WIDE && !wide_ok) return NEFCT_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    double* trig = (double*)workspace;
    hipLaunchKernelGGL(k_trig, dim3((nv + 255) / 256), dim3(256), 0, s, angles, nv, trig);
    if (stageable && mapping != NEFCT_FDK_BP_GATHER) {
        if (mapping == NEFCT_FDK_BP_WIDE) {
            // 1-D grid of every block, rounded up to a multiple of 8 (see XCDMAP)
            const dim3 g((unsigned)((wide_blocks + 7) / 8 * 8));
            hipLaunchKernelGGL((k_fdk_bp_staged<NEFCT_FDK_WIDE>), g, dim3(512), 0, s, filtered, (const double*)trig,
                               offsets, nv, nx, ny, nz, x0, y0, z0, vx, vy, vz, (double)SID, (double)SAD, na,
                               (double)da, (double)ai, nb, (double)db, (double)bi, image);
        } else {
            const dim3 g((nx + 15) / 16, (ny + 15) / 16, (nz + SZ - 1) / SZ);
            hipLaunchKernelGGL((k_fdk_bp_staged<NEFCT_FDK_SMALL>), g, dim3(256), 0, s, filtered, (const double*)trig,
                               offsets, nv, nx, ny, nz, x0, y0, z0, vx, vy, vz, (double)SID, (double)SAD, na,
                               (double)da, (double)ai, nb, (double)db, (double)bi, image);
        }
        NEFCT_CHECK_LAUNCH();
        return NEFCT_OK;
    }
    const dim3 g((nx + 63) / 64, (ny + 3) / 4, (nz + ZC - 1) / ZC);
    if (na >= 2)
        hipLaunchKernelGGL(k_fdk_bp<true>, g, dim3(256), 0, s, filtered, (const double*)trig, offsets, nv, nx, ny, nz,
                           x0, y0, z0, vx, vy, vz, (double)SID, (double)SAD, na, (double)da, (double)ai, nb,
                           (double)db, (double)bi, image);
    else
        hipLaunchKernelGGL(k_fdk_bp<false>, g, dim3(256), 0, s, filtered, (const double*)trig, offsets, nv, nx, ny,
                           nz, x0, y0, z0, vx, vy, vz, (double)SID, (double)SAD, na, (double)da, (double)ai, nb,
                           (double)db, (double)bi, image);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_fdk_backproject(const float* filtered, const int grid[3], const float center[3], const float size[3],
                          const float* angles, const float* offsets, int nv, float SID, float SAD, int na, float da,
                          float ai, int nb, float db, float bi, float* image, void* workspace,
                          size_t workspace_bytes, void* stream)
{
    return nefct_fdk_backproject_mapped(filtered, grid, center, size, angles, offsets, nv, SID, SAD, na, da, ai, nb,
                                        db, bi, image, NEFCT_FDK_BP_AUTO, workspace, workspace_bytes, stream);
}

}  // extern "C"
