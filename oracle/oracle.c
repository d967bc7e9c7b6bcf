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
static void bp_cb(void *p, long v, double l) { bp_ctx *c = (bp_ctx *)p; c->img[v] += c->val * l; }
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
    for (int iv = 0; iv < nv; ++iv)
        for (int ib = 0; ib < G->nb; ++ib)
            for (int ia = 0; ia < G->na; ++ia) {
                double S[3], D[3];
                orc_ray_points(G, angles[iv], ia, ib, S, D);
                bp_ctx c = {img, pv[ia + (long)G->na * (ib + (long)G->nb * iv)]};
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
