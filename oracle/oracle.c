/*
 * oracle.c -- CPU oracle for the cone-beam Siddon projector pair.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file: it is imported by tests/, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg as the timed CPU restatement.
 *
 * Two independent restatements live here:
 *
 *  (1) orc_exact_*  -- FP64 exact Siddon (Siddon 1985, Med. Phys. 12:252;
 *      Jacobs et al. 1998 merged-alpha form).  The ray is the full line
 *      through the source and the detector point, clipped to the volume box;
 *      per-axis plane-crossing parameters are merged in order and every
 *      segment is attributed to the voxel containing its midpoint.  This is
 *      a different algorithm from the slab walk used on the GPU, so a shared
 *      mistake is unlikely.  Forward, exact transpose and an intersection
 *      count (N_int = voxels with nonzero length) are provided.
 *
 *  (2) orc_ref_*  -- FP32 restatement of the reference projector
 *      nefct/cc/siddon_3d/siddon.cu.cc:12-321 (project_device) and
 *      :450-764 (backproject_device), with ray set-up from :341-353 (flat)
 *      and :376-387 (cylindrical).  The reference's quirks are toggles:
 *        ORC_INV_L2     (Q1) weights divided by L^2, L = ray length in x-voxel
 *                       units (siddon.cu.cc:86 et seq.; forward only there)
 *        ORC_REF_SLAB   (Q2) in a slab with both a y and a z crossing the
 *                       reference measures both fractions from the low-value
 *                       end of each axis (:52-76, :132-176); cleared = the
 *                       corrected attribution
 *        ORC_SHORT_SKIP (Q3) adjoint skips rays with in-plane length < 10
 *                       x-voxels (:472)
 *        ORC_TIE_GT     (Q6) adjoint picks the x-slab walk only when
 *                       |xd| > |yd| (:480) instead of >= (:41)
 *      The reference op is unbuildable here (TensorFlow headers and the CUDA
 *      toolkit are absent), so this restatement is pinned by the FP64 oracle
 *      (Q2 cleared => equal to FP32 precision), closed forms and the tests.
 *
 * Geometry (single statement, reference siddon op semantics, siddon.cu.cc:341-353):
 *   source        S = (-SID cos t, -SID sin t, 0)
 *   flat pixel    local (SAD - SID, a, b), cylindrical local (SAD cos a - SID, SAD sin a, b)
 *   a = (ai + ia) da - na da / 2 ,  b = (bi + ib) db - nb db / 2
 *   detector point D = R(t) * local  (rotation about z)
 *   volume: nx*ny*nz voxels centred on (cx,cy,cz), extent (sx,sy,sz) mm,
 *   memory image[ix + nx*(iy + ny*iz)]  (x fastest)
 *   sinogram pv[ia + na*(ib + nb*iv)]
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INV_L2 1u
#define ORC_REF_SLAB 2u
#define ORC_SHORT_SKIP 4u
#define ORC_TIE_GT 8u

typedef struct {
    int nx, ny, nz;
    double cx, cy, cz;
    double sx, sy, sz;
} orc_vol;

typedef struct {
    int cyl;
    double sid, sad; /* reference siddon-op names: SID = source-isocentre, SAD = source-detector */
    int na;
    double da, ai;
    int nb;
    double db, bi;
} orc_det;

/* ------------------------------------------------------------------ */
/* ray end points in double (siddon.cu.cc:341-353, :376-387)          */
/* ------------------------------------------------------------------ */
void orc_ray_points(const orc_det *g, double angle, int ia, int ib, double S[3], double D[3])
{
    const double a = (g->ai + ia) * g->da - g->na * g->da / 2.0;
    const double b = (g->bi + ib) * g->db - g->nb * g->db / 2.0;
    const double c = cos(angle), s = sin(angle);
    double lx, ly;
    if (g->cyl) {
        lx = g->sad * cos(a) - g->sid;
        ly = g->sad * sin(a);
    } else {
        lx = g->sad - g->sid;
        ly = a;
    }
    S[0] = -g->sid * c;
    S[1] = -g->sid * s;
    S[2] = 0.0;
    D[0] = lx * c - ly * s;
    D[1] = lx * s + ly * c;
    D[2] = b;
}

/* ------------------------------------------------------------------ */
/* (1) FP64 exact Siddon, merged alpha                                 */
/* ------------------------------------------------------------------ */
typedef void (*orc_seg_fn)(void *ctx, long vox, double len_mm);

/* Walks every voxel the line S->D (extended) crosses inside the box; calls
 * fn(ctx, voxel, length_mm) for each segment of positive length. */
static void exact_walk(const orc_vol *V, const double S[3], const double D[3], orc_seg_fn fn, void *ctx)
{
    const int n[3] = {V->nx, V->ny, V->nz};
    const double c[3] = {V->cx, V->cy, V->cz};
    const double sz[3] = {V->sx, V->sy, V->sz};
    double lo[3], vs[3], d[3];
    double amin = -INFINITY, amax = INFINITY;
    for (int i = 0; i < 3; ++i) {
        vs[i] = sz[i] / n[i];
        lo[i] = c[i] - sz[i] / 2.0;
        d[i] = D[i] - S[i];
        const double hi = lo[i] + sz[i];
        if (d[i] == 0.0) {
            if (!(S[i] > lo[i] && S[i] < hi)) return;
            continue;
        }
        double a0 = (lo[i] - S[i]) / d[i], a1 = (hi - S[i]) / d[i];
        if (a0 > a1) { double t = a0; a0 = a1; a1 = t; }
        if (a0 > amin) amin = a0;
        if (a1 < amax) amax = a1;
    }
    if (!(amax > amin)) return;
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    /* per-axis next crossing index and step */
    int k[3], step[3], kend[3];
    for (int i = 0; i < 3; ++i) {
        if (d[i] == 0.0) { k[i] = 0; step[i] = 0; kend[i] = 0; continue; }
        const double pmin = (S[i] + amin * d[i] - lo[i]) / vs[i];
        const double pmax = (S[i] + amax * d[i] - lo[i]) / vs[i];
        if (d[i] > 0) { k[i] = (int)floor(pmin) + 1; step[i] = 1; kend[i] = (int)ceil(pmax) - 1; }
        else { k[i] = (int)ceil(pmin) - 1; step[i] = -1; kend[i] = (int)floor(pmax) + 1; }
    }
    double acur = amin;
    for (;;) {
        double anext = amax;
        for (int i = 0; i < 3; ++i) {
            if (!step[i]) continue;
            if ((step[i] > 0 && k[i] > kend[i]) || (step[i] < 0 && k[i] < kend[i])) continue;
            const double a = (lo[i] + k[i] * vs[i] - S[i]) / d[i];
            if (a < anext) anext = a;
        }
        if (anext > acur) {
            const double am = 0.5 * (acur + anext);
            long idx[3];
            int ok = 1;
            for (int i = 0; i < 3; ++i) {
                idx[i] = (long)floor((S[i] + am * d[i] - lo[i]) / vs[i]);
                if (idx[i] < 0) idx[i] = 0;
                if (idx[i] >= n[i]) idx[i] = n[i] - 1;
            }
            if (ok) fn(ctx, idx[0] + (long)n[0] * (idx[1] + (long)n[1] * idx[2]), (anext - acur) * len);
        }
        if (anext >= amax) break;
        /* advance every axis whose crossing equals anext */
        for (int i = 0; i < 3; ++i) {
            if (!step[i]) continue;
            if ((step[i] > 0 && k[i] > kend[i]) || (step[i] < 0 && k[i] < kend[i])) continue;
            const double a = (lo[i] + k[i] * vs[i] - S[i]) / d[i];
            if (a <= anext) k[i] += step[i];
        }
        acur = anext;
    }
}

typedef struct { const double *img; double acc; } fwd_ctx;
static void fwd_cb(void *p, long v, double l) { fwd_ctx *c = (fwd_ctx *)p; c->acc += c->img[v] * l; }
typedef struct { double *img; double val; } bp_ctx;
static void bp_cb(void *p, long v, double l)
{
    bp_ctx *c = (bp_ctx *)p;
    const double add = c->val * l;
#pragma omp atomic
    c->img[v] += add;
}
typedef struct { long count; double sum_len; } cnt_ctx;
static void cnt_cb(void *p, long v, double l) { (void)v; cnt_ctx *c = (cnt_ctx *)p; c->count += 1; c->sum_len += l; }

/* Line integral of one ray given explicit end points (mm). */
double orc_exact_ray(const orc_vol *V, const double *img, const double S[3], const double D[3])
{
    fwd_ctx c = {img, 0.0};
    exact_walk(V, S, D, fwd_cb, &c);
    return c.acc;
}

/* Chord length of one ray with the volume box and its intersection count. */
long orc_exact_ray_count(const orc_vol *V, const double S[3], const double D[3], double *sum_len)
{
    cnt_ctx c = {0, 0.0};
    exact_walk(V, S, D, cnt_cb, &c);
    if (sum_len) *sum_len = c.sum_len;
    return c.count;
}

/* Forward projection of the whole sinogram, FP64 (out[nv*nb*na]). */
void orc_exact_fwd(const orc_vol *V, const orc_det *G, const double *angles, int nv,
                   const double *img, double *out)
{
    const long nrays = (long)nv * G->nb * G->na;
#pragma omp parallel for schedule(dynamic, 64)
    for (long r = 0; r < nrays; ++r) {
        const int ia = (int)(r % G->na), ib = (int)((r / G->na) % G->nb), iv = (int)(r / ((long)G->na * G->nb));
        double S[3], D[3];
        orc_ray_points(G, angles[iv], ia, ib, S, D);
        out[r] = orc_exact_ray(V, img, S, D);
    }
}

/* Exact transpose of orc_exact_fwd (img zeroed here). */
void orc_exact_bp(const orc_vol *V, const orc_det *G, const double *angles, int nv,
                  const double *pv, double *img)
{
    const long nvox = (long)V->nx * V->ny * V->nz;
    memset(img, 0, sizeof(double) * nvox);
    /* rays in parallel, each voxel update atomic: the summation order varies
     * between runs at the level of FP64 rounding (~1e-16 relative) */
    const long nrays = (long)nv * G->nb * G->na;
#pragma omp parallel for schedule(dynamic, 64)
    for (long r = 0; r < nrays; ++r) {
        const int ia = (int)(r % G->na), ib = (int)((r / G->na) % G->nb), iv = (int)(r / ((long)G->na * G->nb));
        double S[3], D[3];
        orc_ray_points(G, angles[iv], ia, ib, S, D);
        bp_ctx c = {img, pv[r]};
        exact_walk(V, S, D, bp_cb, &c);
    }
}

/* Per-ray intersection counts (voxels with nonzero length), and the total. */
long orc_exact_count(const orc_vol *V, const orc_det *G, const double *angles, int nv, int32_t *counts)
{
    const long nrays = (long)nv * G->nb * G->na;
    long total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (long r = 0; r < nrays; ++r) {
        const int ia = (int)(r % G->na), ib = (int)((r / G->na) % G->nb), iv = (int)(r / ((long)G->na * G->nb));
        double S[3], D[3];
        orc_ray_points(G, angles[iv], ia, ib, S, D);
        const long c = orc_exact_ray_count(V, S, D, NULL);
        if (counts) counts[r] = (int32_t)c;
        total += c;
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* (2) restatement of the reference slab walk, FP32 and FP64           */
/* ------------------------------------------------------------------ */
#define REAL float
#define SFX(n) n##_f32
#define FLOOR floorf
#define SQRT sqrtf
#define FABS fabsf
#define COS cosf
#define SIN sinf
#include "ref_walk.h"
#undef REAL
#undef SFX
#undef FLOOR
#undef SQRT
#undef FABS
#undef COS
#undef SIN
#define REAL double
#define SFX(n) n##_f64
#define FLOOR floor
#define SQRT sqrt
#define FABS fabs
#define COS cos
#define SIN sin
#include "ref_walk.h"

/* the FP32 entry points keep their plain names */
void orc_ref_fwd(const orc_vol *V, const orc_det *G, const float *angles, int nv,
                 const float *img, float *out, unsigned flags, int v0, int vstep)
{
    orc_ref_fwd_f32(V, G, angles, nv, img, out, flags, v0, vstep);
}

void orc_ref_bp(const orc_vol *V, const orc_det *G, const float *angles, int nv,
                const float *pv, float *out, unsigned flags, int v0, int vstep)
{
    orc_ref_bp_f32(V, G, angles, nv, pv, out, flags, v0, vstep);
}

/* Threads the oracle would use (for the baseline record). */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------
 * Voxel-driven backprojection without weights, restated in FP64 from
 * nefct/cc/pixel_3d/pixel.cu.cc:12-124 (flat :12-52, cylindrical :54-95,
 * voxel centres :116-118).  The reference's integer n/2 in the detector
 * index is kept.  image [nz][ny][nx] is overwritten.
 * ------------------------------------------------------------------------ */
void orc_pixel_bp(int cyl, const int grid[3], const double center[3], const double size[3],
                  const double *angles, int nv, double SID, double SAD, int na, double da, double ai,
                  int nb, double db, double bi, const double *pv, double *image)
{
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    const double ha = (double)(na / 2), hb = (double)(nb / 2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const double xc = (ix + 0.5) * size[0] / nx + center[0];
                const double yc = (iy + 0.5) * size[1] / ny + center[1];
                const double zc = (iz + 0.5) * size[2] / nz + center[2];
                double acc = 0.0;
                for (int iv = 0; iv < nv; ++iv) {
                    const double c = cos(angles[iv]), s = sin(angles[iv]);
                    const double x0 = xc * c + yc * s;
                    const double y0 = -xc * s + yc * c;
                    const double mag = SAD / (x0 + SID);
                    const double a = cyl ? atan2(y0, x0 + SID) / da + ha - ai - 0.5
                                         : mag * y0 / da + ha - ai - 0.5;
                    const double b = mag * zc / db + hb - bi - 0.5;
                    const int ia = (int)floor(a), ib = (int)floor(b);
                    if (ia < 0 || ia + 1 >= na || ib < 0 || ib + 1 >= nb) continue;
                    const double wa2 = a - ia, wa = 1.0 - wa2, wb2 = b - ib, wb = 1.0 - wb2;
                    const double *p = pv + (size_t)iv * na * nb;
                    acc += p[ia + (size_t)ib * na] * wa * wb + p[ia + (size_t)(ib + 1) * na] * wa * wb2 +
                           p[ia + 1 + (size_t)ib * na] * wa2 * wb + p[ia + 1 + (size_t)(ib + 1) * na] * wa2 * wb2;
                }
                image[ix + (size_t)nx * (iy + (size_t)ny * iz)] = acc;
            }
}

/* ------------------------------------------------------------------------
 * Pixel-driven FDK backprojection with slab footprints, restated in FP64
 * from nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255.  Lengths are in
 * voxel units with the volume centred on the axis (the reference's Python
 * layer divides by the voxel size); that file names the source-to-axis
 * distance SAD and the source-to-detector distance SID.  For every pixel
 * (ia, ib, iv) and every slab of its dominant in-plane axis, the pixel's
 * footprint on the slab's mid-plane is bounded by its two edge rays in the
 * transverse in-plane axis (all four slope combinations, :104-125) and by
 * the centre ray +- db/2 in z (:127-138); each overlapped voxel receives
 * pv * wy * wz * |r| / |r_dom| * SAD^2 / (SAD + x cos + y sin)^2 (:142-157).
 * offsets[iv] (NULL = 0) raises source and detector along z.  The integer
 * divisions n/2 of the reference are kept where it has them.  image
 * [nz][ny][nx] is overwritten.
 * ------------------------------------------------------------------------ */
static void fdk_fp_edge(int cyl, double SID, double SAD, double da, double ai, int na, double e, double c,
                        double s, double x1, double y1, double *xo, double *yo)
{
    double x20, y20;
    if (!cyl) {
        x20 = SID - SAD;
        y20 = e * da + ai - na * da / 2;
    } else {
        const double ang = e * da + ai - na * da / 2;
        x20 = SID * cos(ang) - SAD;
        y20 = SID * sin(ang);
    }
    *xo = x20 * c - y20 * s - x1;
    *yo = x20 * s + y20 * c - y1;
}

void orc_fdk_fp_bp(int cyl, int nx, int ny, int nz, const double *angles, const double *offsets, int nv,
                   double SID, double SAD, int na, double da, double ai, int nb, double db, double bi,
                   const double *pv, double *image)
{
    const size_t nvox = (size_t)nx * ny * nz;
    memset(image, 0, sizeof(double) * nvox);
    const int hx = nx / 2, hy = ny / 2, hz = nz / 2;
#pragma omp parallel
    {
        double *priv = (double *)calloc(nvox, sizeof(double));
#pragma omp for collapse(2) schedule(dynamic, 1)
        for (int iv = 0; iv < nv; ++iv)
            for (int ib = 0; ib < nb; ++ib)
                for (int ia = 0; ia < na; ++ia) {
                    const double c = cos(angles[iv]), s = sin(angles[iv]);
                    const double off = offsets ? offsets[iv] : 0.0;
                    const double x1 = -SAD * c, y1 = -SAD * s, z1 = off;
                    double x20, y20;
                    if (!cyl) {
                        x20 = SID - SAD;
                        y20 = (ia + 0.5) * da + ai - na * da / 2;
                    } else {
                        const double ang = (ia + 0.5) * da + ai - na * da / 2;
                        x20 = SID * cos(ang) - SAD;
                        y20 = SID * sin(ang);
                    }
                    const double z20 = (ib + 0.5) * db + bi - nb * db / 2;
                    const double x21 = x20 * c - y20 * s - x1;
                    const double y21 = x20 * s + y20 * c - y1;
                    const double z21 = z20 + off - z1;
                    const double val = pv[ia + (size_t)na * (ib + (size_t)nb * iv)];
                    const double len = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
                    double xl, yl, xr, yr;
                    fdk_fp_edge(cyl, SID, SAD, da, ai, na, (double)ia, c, s, x1, y1, &xl, &yl);
                    fdk_fp_edge(cyl, SID, SAD, da, ai, na, (double)(ia + 1), c, s, x1, y1, &xr, &yr);
                    const int xdom = fabs(x21) > fabs(y21);
                    const int nm = xdom ? nx : ny;      /* walk axis */
                    for (int im = 0; im < nm; ++im) {
                        /* distance of the slab mid-plane from the source along the walk axis */
                        const double mc = xdom ? (im + 0.5 - nx / 2.0 - x1) : (im + 0.5 - ny / 2.0 - y1);
                        double q[4];
                        if (xdom) {
                            q[0] = yl / xl * mc; q[1] = yl / xr * mc; q[2] = yr / xl * mc; q[3] = yr / xr * mc;
                        } else {
                            q[0] = xl / yl * mc; q[1] = xl / yr * mc; q[2] = xr / yl * mc; q[3] = xr / yr * mc;
                        }
                        const double base_t = xdom ? (y1 + hy) : (x1 + hx);
                        double t1 = q[0] + base_t, t2 = t1;
                        for (int k = 1; k < 4; ++k) {
                            const double v = q[k] + base_t;
                            if (v < t1) t1 = v;
                            if (v > t2) t2 = v;
                        }
                        const double rdom = xdom ? x21 : y21;
                        const double zl = (z21 - db / 2) / rdom * mc + z1 + hz;
                        const double zr = (z21 + db / 2) / rdom * mc + z1 + hz;
                        const double z1f = zl < zr ? zl : zr, z2f = zl < zr ? zr : zl;
                        const int T1 = (int)floor(t1), T2 = (int)floor(t2);
                        const int Z1 = (int)floor(z1f), Z2 = (int)floor(z2f);
                        const int nt = xdom ? ny : nx;
                        for (int it = (T1 > 0 ? T1 : 0); it <= (T2 < nt - 1 ? T2 : nt - 1); ++it) {
                            const double wt = ((it + 1.0 < t2 ? it + 1.0 : t2) - (it > t1 ? it : t1)) / (t2 - t1);
                            for (int iz = (Z1 > 0 ? Z1 : 0); iz <= (Z2 < nz - 1 ? Z2 : nz - 1); ++iz) {
                                const double wz = ((iz + 1.0 < z2f ? iz + 1.0 : z2f) - (iz > z1f ? iz : z1f)) /
                                                  (z2f - z1f);
                                const int ix = xdom ? im : it, iy = xdom ? it : im;
                                const double xc = ix + 0.5 - hx, yc = iy + 0.5 - hy;
                                const double u = SAD + xc * c + yc * s;
                                const double w = wt * wz / fabs(rdom) * len * SAD * SAD / u / u;
                                priv[ix + (size_t)nx * (iy + (size_t)ny * iz)] += val * w;
                            }
                        }
                    }
                }
#pragma omp critical
        for (size_t i = 0; i < nvox; ++i) image[i] += priv[i];
        free(priv);
    }
}
