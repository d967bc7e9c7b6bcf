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
// detector row, a workgroup 4 such rows; the workgroup's additions are summed
// in LDS over short chunks of slabs before they reach the volume (kernel
// comment below).  The per-pixel edge rays, slopes, |r| and z bounds are
// computed once per pixel (the reference recomputes the edge rays, with
// cosf/sinf for the cylinder, in every slab).  Volume additions go through
// buffer atomics whose range check covers the volume.
#include <math.h>

#include <cmath>
#include <initializer_list>
#include <type_traits>

#include "common.h"

namespace {

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

// floor(v) clamped to [lo, hi] by one med3 (a NaN converts to 0, inside
// every range used here)
__device__ __forceinline__ int floor_med(float v, float lo, float hi)
{
    return (int)__builtin_amdgcn_fmed3f(floorf(v), lo, hi);
}

// In-plane components of the ray from the source to detector coordinate e.
template <bool CYL>
__device__ __forceinline__ void ray_to(float SID, float SAD, float da, float ai, int na, float e, float c, float s,
                                       float& x, float& y)
{
    edge_point<CYL>(SID, SAD, da, ai, na, e, c, s, -SAD * c, -SAD * s, x, y);
}

// Per-pixel quantities of one view for walk axis m (x when XDOM) and
// transverse in-plane axis t, computed once per pixel; the reference
// recomputes the edge rays in every slab.  The four edge-ray pairings'
// transverse positions q_k * mc + base_t all meet at the source (mc = 0), so
// their minimum is the one with the smallest slope when mc >= 0 and the
// largest when mc < 0 (fma rounding is monotone in q_k): two fmas give the
// same hull as four fmas and six min/max.  Likewise for the two z slopes.
struct Pix {
    float qlo, qhi;       // smallest / largest transverse slope
    float src_m, base_t;  // source on the walk axis, transverse offset
    float kzlo, kzhi, base_z;
    float half_m, vpre;   // nm / 2, |r| / |r_dom| * SAD^2
    float rdom;
};

template <bool CYL, bool XDOM>
__device__ __forceinline__ Pix pixel_setup(int ia, int ib, float c, float s, float off, float SID, float SAD,
                                           int nx, int ny, int nz, float da, float ai, int na, float db, float bi,
                                           int nb)
{
    Pix p;
    const float x1 = -SAD * c, y1 = -SAD * s, z1 = off;
    float x21, y21, xl, yl, xr, yr;
    ray_to<CYL>(SID, SAD, da, ai, na, (float)ia + 0.5f, c, s, x21, y21);
    ray_to<CYL>(SID, SAD, da, ai, na, (float)ia, c, s, xl, yl);
    ray_to<CYL>(SID, SAD, da, ai, na, (float)(ia + 1), c, s, xr, yr);
    const float z21 = ((ib + 0.5f) * db + bi - nb * db / 2 + off) - z1;
    const float len = sqrtf(x21 * x21 + y21 * y21 + z21 * z21);
    const float rdom = XDOM ? x21 : y21;
    float q0, q1, q2, q3;
    if (XDOM) {
        q0 = yl / xl; q1 = yl / xr; q2 = yr / xl; q3 = yr / xr;
    } else {
        q0 = xl / yl; q1 = xl / yr; q2 = xr / yl; q3 = xr / yr;
    }
    p.qlo = fminf(fminf(q0, q1), fminf(q2, q3));
    p.qhi = fmaxf(fmaxf(q0, q1), fmaxf(q2, q3));
    p.src_m = XDOM ? x1 : y1;
    p.base_t = (XDOM ? y1 : x1) + (float)((XDOM ? ny : nx) / 2);
    const float kzl = (z21 - db / 2) / rdom, kzr = (z21 + db / 2) / rdom;
    p.kzlo = fminf(kzl, kzr);
    p.kzhi = fmaxf(kzl, kzr);
    p.base_z = z1 + (float)(nz / 2);
    p.half_m = (float)(XDOM ? nx : ny) / 2;
    p.vpre = len / fabsf(rdom) * SAD * SAD;
    p.rdom = rdom;
    return p;
}

__device__ __forceinline__ float walk_coord(const Pix& p, int im)
{
    return (float)im + 0.5f - p.half_m - p.src_m;
}

// Transverse and z extent of the pixel's footprint at walk coordinate mc.
__device__ __forceinline__ void hull(const Pix& p, float mc, float& t1, float& t2, float& z1f, float& z2f)
{
    const bool fwd = mc >= 0.0f;
    t1 = (fwd ? p.qlo : p.qhi) * mc + p.base_t;
    t2 = (fwd ? p.qhi : p.qlo) * mc + p.base_t;
    z1f = (fwd ? p.kzlo : p.kzhi) * mc + p.base_z;
    z2f = (fwd ? p.kzhi : p.kzlo) * mc + p.base_z;
}

// One slab of one pixel: every overlapped voxel gets
// val * wt * wz * pre / U^2, handed to put(it, iz, value) in walk-axis
// coordinates (voxel x = im, y = it when XDOM; x = it, y = im otherwise).
template <bool XDOM, class Put>
__device__ __forceinline__ void slab(const Pix& p, float val, int im, int nt, int nz, int nx, int ny, float c,
                                     float s, float SAD, Put&& put)
{
    float t1, t2, z1f, z2f;
    hull(p, walk_coord(p, im), t1, t2, z1f, z2f);
    const int it0 = floor_med(t1, 0.0f, (float)nt), it1 = floor_med(t2, -1.0f, (float)(nt - 1));
    const int iz0 = floor_med(z1f, 0.0f, (float)nz), iz1 = floor_med(z2f, -1.0f, (float)(nz - 1));
    if (it0 > it1 || iz0 > iz1) return;
    const int hx = nx / 2, hy = ny / 2;
    const float inv_t = 1.0f / (t2 - t1), inv_z = 1.0f / (z2f - z1f);
    const float pre = p.vpre;
    for (int it = it0; it <= it1; ++it) {
        const float wt = (fminf(it + 1.0f, t2) - fmaxf((float)it, t1)) * inv_t;
        const int ix = XDOM ? im : it, iy = XDOM ? it : im;
        const float xc = ix + 0.5f - hx, yc = iy + 0.5f - hy;
        const float u = SAD + xc * c + yc * s;
        const float w_t = val * wt * pre / (u * u);
        for (int iz = iz0; iz <= iz1; ++iz) {
            const float wz = (fminf(iz + 1.0f, z2f) - fmaxf((float)iz, z1f)) * inv_z;
            put(it, iz, w_t * wz);
        }
    }
}

// A workgroup = 64 neighbouring columns x 4 rows of one view.  Its pixels'
// footprints on S consecutive slabs cover a small (slab, t, z) box, so the
// slabs are walked in chunks of S: each pixel adds into an LDS window over the
// chunk's box, then the window is added into the volume with one global
// atomic per touched voxel, neighbouring lanes on neighbouring voxels of the
// volume's contiguous axis.  Neighbouring pixels overlap the same voxels (a
// 0.8 mm voxel seen through 0.6 mm columns at magnification 1.5 spans two
// columns), so this cuts the global atomics several-fold and gathers them into
// whole cache lines, where direct per-pixel atomics along x hit one line per
// lane.  The window holds 64-bit fixed point (ds_add_u64, each contribution
// scaled by the workgroup's power of two 2^e with |contribution| 2^e <= 2^30
// and truncated; the rounding step is ~1e-9 of the workgroup's largest
// contribution): integer LDS atomics run several times faster than float ones
// (ds_add_f32) on gfx950.  At 512 x 512 x 256 voxels and 1024 x 768 pixels:
// 24 ms per view with direct atomics, 3.6 with a float window, 2.0 with this.
//
// The box is computed by every wavefront alike from the workgroup's outer
// pixels (detector positions, and hence footprints, are monotone in the
// column and row), with one cell of slack; a deposit outside it still goes
// straight to the volume.  Two windows alternate, so one barrier per chunk
// separates a chunk's deposits from its flush and the previous flush of the
// same window from the next deposits.  A chunk whose box exceeds the window,
// a workgroup whose pixels disagree on the walk axis, and one with a
// non-finite value or voxels within 1e-3 SAD of the source's plane add
// directly (float atomics into the volume).
//
// Window layout [z][m][t] with a row pitch of WT + 1 words: the lanes of a
// deposit (neighbouring t at one slab) and of a flush along m (an x-walk's
// contiguous axis) both fall in distinct banks.
template <bool CYL, int S, int WT, int WZ>
__global__ __launch_bounds__(256) void k_fdk_fp_bp(const float* __restrict__ pv, const float* __restrict__ angles,
                                                   const float* __restrict__ offsets, int nv, float SID, float SAD,
                                                   int nx, int ny, int nz, float da, float ai, int na, float db,
                                                   float bi, int nb, float* __restrict__ image)
{
    static_assert((S & (S - 1)) == 0, "chunk length");
    constexpr int WP = WT + 1;             // row pitch: rows start in successive banks
    constexpr int WIN = S * WP * WZ;
    __shared__ unsigned long long win[2 * WIN];
    __shared__ unsigned s_cmax;
    const int lane = threadIdx.x & 63;
    const int ia = blockIdx.x * 64 + lane;
    const int ib0 = blockIdx.y * 4, ib = ib0 + (threadIdx.x >> 6);
    const int iv = blockIdx.z;
    const bool on = ia < na && ib < nb;
    const float val = on ? pv[(size_t)ia + (size_t)na * ((size_t)ib + (size_t)nb * iv)] : 0.0f;
    const bool act = on && val != 0.0f;   // a zero pixel adds nothing
    const double ang = (double)angles[iv];
    const float c = (float)cos(ang), s = (float)sin(ang);
    const float off = offsets ? offsets[iv] : 0.0f;
    float x21, y21;
    ray_to<CYL>(SID, SAD, da, ai, na, (float)ia + 0.5f, c, s, x21, y21);
    const bool xdom = fabsf(x21) > fabsf(y21);
    const __amdgpu_buffer_rsrc_t img = make_rsrc(image, (uint32_t)((size_t)nx * ny * nz * 4));
    auto add = [&](int ix, int iy, int iz, float v) { buf_atomic_add(img, (ix + nx * (iy + ny * iz)) * 4, v); };

    if (threadIdx.x == 0) s_cmax = 0u;
    const int any = __syncthreads_or(act);
    if (!any) return;
    const int all_x = __syncthreads_and(!act || xdom);
    const int all_y = __syncthreads_and(!act || !xdom);

    auto run = [&](auto xdom_tag, auto uniform_tag) {
        constexpr bool WX = decltype(xdom_tag)::value;
        constexpr bool UNI = decltype(uniform_tag)::value;
        const Pix p = pixel_setup<CYL, WX>(ia, ib, c, s, off, SID, SAD, nx, ny, nz, da, ai, na, db, bi, nb);
        const int NM = WX ? nx : ny, NT = WX ? ny : nx;
        auto direct = [&](int im) {
            slab<WX>(p, val, im, NT, nz, nx, ny, c, s, SAD,
                     [&](int it, int iz, float v) { WX ? add(im, it, iz, v) : add(it, im, iz, v); });
        };
        if constexpr (!UNI) {   // the walk axis changes inside the workgroup
            if (act)
                for (int im = 0; im < NM; ++im) direct(im);
            return;
        }
        // fixed-point scale 2^e of the workgroup: every contribution is at
        // most max |pv| pre / umin^2 (overlap fractions <= 1; umin the
        // smallest U = SAD + x cos + y sin over the voxel centres)
        if (act) atomicMax(&s_cmax, __float_as_uint(fabsf(val) * p.vpre));   // >= 0: float order = bit order
        for (int i = threadIdx.x; i < 2 * WIN; i += 256) win[i] = 0ull;
        __syncthreads();
        float umin = 3.0e38f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float xq = (q & 1) ? nx - 0.5f - nx / 2 : 0.5f - nx / 2;
            const float yq = (q & 2) ? ny - 0.5f - ny / 2 : 0.5f - ny / 2;
            umin = fminf(umin, SAD + xq * c + yq * s);
        }
        const double bound = (double)__uint_as_float(s_cmax) / ((double)umin * (double)umin);
        // non-finite data or a volume reaching the source: direct additions
        const bool fixed = s_cmax <= __float_as_uint(3.0e38f) && umin > 1e-3f * SAD && bound > 0.0;
        int e = 0;
        if (fixed) {
            frexp(1073741824.0 / bound, &e);   // 2^30 / bound = m 2^e, m in [0.5, 1)
            e -= 1;
        }
        // the workgroup's outer pixels: columns of lanes 0 and 63 (every
        // wavefront holds the same columns), rows ib0 and ib0 + 3
        const int last = min(63, na - 1 - (int)blockIdx.x * 64);
        auto lane_val = [](float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); };
        const float qlo = fminf(lane_val(p.qlo, 0), lane_val(p.qlo, last));
        const float qhi = fmaxf(lane_val(p.qhi, 0), lane_val(p.qhi, last));
        const float rd0 = lane_val(p.rdom, 0), rd1 = lane_val(p.rdom, last);
        const float zb = (ib0 + 0.5f) * db + bi - nb * db / 2 - db / 2;
        const float zt = (min(ib0 + 3, nb - 1) + 0.5f) * db + bi - nb * db / 2 + db / 2;
        const float kz[4] = {zb / rd0, zb / rd1, zt / rd0, zt / rd1};
        float kzlo = kz[0], kzhi = kz[0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            kzlo = fminf(kzlo, kz[k]);
            kzhi = fmaxf(kzhi, kz[k]);
        }
        int buf = 0;
        for (int im0 = 0; im0 < NM; im0 += S) {
            const int se = min(S, NM - im0);
            // hull and z extents are linear in the walk coordinate on either
            // side of the source, so the box's extremes lie on the chunk's ends
            float tlo = 3.0e38f, thi = -3.0e38f, zlo = 3.0e38f, zhi = -3.0e38f;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float mc = walk_coord(p, e ? im0 + se - 1 : im0);
                const bool fwd = mc >= 0.0f;
                tlo = fminf(tlo, (fwd ? qlo : qhi) * mc + p.base_t);
                thi = fmaxf(thi, (fwd ? qhi : qlo) * mc + p.base_t);
                zlo = fminf(zlo, (fwd ? kzlo : kzhi) * mc + p.base_z);
                zhi = fmaxf(zhi, (fwd ? kzhi : kzlo) * mc + p.base_z);
            }
            const int T0 = max(floor_med(tlo, 0.0f, (float)NT) - 1, 0);
            const int T1 = min(floor_med(thi, -1.0f, (float)(NT - 1)) + 1, NT - 1);
            const int Z0 = max(floor_med(zlo, 0.0f, (float)nz) - 1, 0);
            const int Z1 = min(floor_med(zhi, -1.0f, (float)(nz - 1)) + 1, nz - 1);
            if (T0 > T1 || Z0 > Z1) continue;   // no pixel meets the volume here
            const int ET = T1 - T0 + 1, EZ = Z1 - Z0 + 1;
            if (!fixed || ET > WT || EZ > WZ) {
                if (act)
                    for (int im = im0; im < im0 + se; ++im) direct(im);
                continue;
            }
            unsigned long long* w = win + buf * WIN;
            buf ^= 1;
            if (act) {
                for (int im = im0; im < im0 + se; ++im) {
                    const int lm = im - im0;
                    slab<WX>(p, val, im, NT, nz, nx, ny, c, s, SAD, [&](int it, int iz, float v) {
                        const int lt = it - T0, lz = iz - Z0;
                        if ((unsigned)lt < (unsigned)ET && (unsigned)lz < (unsigned)EZ) {
                            // |v 2^e| <= 2^30; truncated toward zero, sign-extended
                            const int qv = (int)ldexpf(v, e);
                            atomicAdd(&w[(lz * S + lm) * WP + lt], (unsigned long long)(long long)qv);
                        } else {
                            WX ? add(im, it, iz, v) : add(it, im, iz, v);
                        }
                    });
                }
            }
            __syncthreads();
            // flush the box, EZ x ET x S cells, along the volume's x (m for
            // an x-walk, t for a y-walk); r / ET by a float reciprocal
            // (r + 0.5 stays >= 0.5 / ET away from a multiple)
            const float inv_et = 1.0f / (float)ET;
            const int ncell = EZ * ET * S;
            for (int cidx = threadIdx.x; cidx < ncell; cidx += 256) {
                int lm, lt, lz;
                if (WX) {
                    lm = cidx & (S - 1);
                    const int r = cidx / S;
                    lz = (int)(((float)r + 0.5f) * inv_et);
                    lt = r - lz * ET;
                } else {
                    const int r = (int)(((float)cidx + 0.5f) * inv_et);
                    lt = cidx - r * ET;
                    lz = r / S;
                    lm = r & (S - 1);
                }
                const int k = (lz * S + lm) * WP + lt;
                const int iz = Z0 + lz, ix = WX ? im0 + lm : T0 + lt, iy = WX ? T0 + lt : im0 + lm;
                const unsigned long long v = w[k];
                if (v == 0ull) continue;
                w[k] = 0ull;
                add(ix, iy, iz, (float)ldexp((double)(long long)v, -e));
            }
        }
    };
    if (all_x)
        run(std::true_type{}, std::true_type{});
    else if (all_y)
        run(std::false_type{}, std::true_type{});
    else if (xdom)
        run(std::true_type{}, std::false_type{});
    else
        run(std::false_type{}, std::false_type{});
}

// chunks of 4 slabs, windows of 64 transverse x 8 z cells (2 x 16.6 KB of LDS)
constexpr int FS = 4, FWT = 64, FWZ = 8;

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
        hipLaunchKernelGGL((k_fdk_fp_bp<true, FS, FWT, FWZ>), g, dim3(256), 0, s, pv, angles, offsets, nv, SID, SAD, nx, ny, nz, da,
                           ai, na, db, bi, nb, image);
    else
        hipLaunchKernelGGL((k_fdk_fp_bp<false, FS, FWT, FWZ>), g, dim3(256), 0, s, pv, angles, offsets, nv, SID, SAD, nx, ny, nz, da,
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
