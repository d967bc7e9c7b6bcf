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
/* (2) FP32 restatement of the reference slab walk                     */
/* ------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz;
    float cx, cy, cz, sx, sy, sz;
} vol32;

/* Reference ray set-up in float (siddon.cu.cc:341-353 / :376-387). */
static void ref_points32(const orc_det *g, float angle, int ia, int ib, float S[3], float D[3])
{
    const float na = (float)g->na, nb = (float)g->nb;
    const float da = (float)g->da, db = (float)g->db, ai = (float)g->ai, bi = (float)g->bi;
    const float SID = (float)g->sid, SAD = (float)g->sad;
    const float a = (ai + ia) * da - na * da / 2;
    const float b = (bi + ib) * db - nb * db / 2;
    const float c = cosf(angle), s = sinf(angle);
    float lx, ly;
    if (g->cyl) {
        lx = SAD * cosf(a) - SID;
        ly = SAD * sinf(a);
    } else {
        lx = SAD - SID;
        ly = a;
    }
    S[0] = -SID * c;
    S[1] = -SID * s;
    S[2] = 0.0f;
    D[0] = lx * c - ly * s;
    D[1] = lx * s + ly * c;
    D[2] = b;
}

/* Visitor for one weighted voxel: forward accumulates, adjoint scatters. */
typedef struct {
    const float *img; /* forward */
    float *out;       /* adjoint */
    float val;        /* adjoint ray value */
    float acc;        /* forward accumulator */
    int fwd;
} ref_acc;

static inline void ref_visit(ref_acc *A, long idx, float w)
{
    if (A->fwd) A->acc += A->img[idx] * w;
    else A->out[idx] += A->val * w;
}

/*
 * One slab of the reference walk.  'lo'/'hi' are the sorted transverse
 * coordinates (index units) at the two slab faces, as the reference computes
 * yy1<=yy2 and zz1<=zz2 (siddon.cu.cc:52-76).  'flip_z' is true when the z
 * ordering is measured from the opposite slab face to the y ordering
 * (sign(ky) != sign(kz)); the reference ignores it (Q2), the corrected rule
 * uses it.
 */
static void ref_slab(ref_acc *A, float p1, float p2, float q1, float q2, int np, int nq,
                     long base, long sp, long sq, float wbase, int ref_slab_rule, int flip_q)
{
    const int cp1 = (int)floorf(p1), cp2 = (int)floorf(p2);
    const int cq1 = (int)floorf(q1), cq2 = (int)floorf(q2);
#define VOX(ip, iq) (base + (long)(ip) * sp + (long)(iq) * sq)
    if (cp1 == cp2) {
        if (!(0 <= cp1 && cp1 < np)) return;
        if (cq1 == cq2) {
            if (0 <= cq1 && cq1 < nq) ref_visit(A, VOX(cp1, cq1), wbase);
        } else if (-1 <= cq1 && cq1 < nq) {
            const float rq = (cq2 - q1) / (q2 - q1);
            if (cq1 >= 0) ref_visit(A, VOX(cp1, cq1), rq * wbase);
            if (cq2 < nq) ref_visit(A, VOX(cp1, cq2), (1 - rq) * wbase);
        }
        return;
    }
    if (!(-1 <= cp1 && cp1 < np)) return;
    if (cq1 == cq2) {
        if (0 <= cq1 && cq1 < nq) {
            const float rp = (cp2 - p1) / (p2 - p1);
            if (cp1 >= 0) ref_visit(A, VOX(cp1, cq1), rp * wbase);
            if (cp2 < np) ref_visit(A, VOX(cp2, cq1), (1 - rp) * wbase);
        }
        return;
    }
    if (!(-1 <= cq1 && cq1 < nq)) return;
    const float rp = (cp2 - p1) / (p2 - p1);
    const float rq = (cq2 - q1) / (q2 - q1);
    if (ref_slab_rule || !flip_q) {
        /* reference attribution: both fractions from the low end (siddon.cu.cc:132-176) */
        if (rp > rq) {
            if (cp1 >= 0 && cq1 >= 0) ref_visit(A, VOX(cp1, cq1), rq * wbase);
            if (cp1 >= 0 && cq2 < nq) ref_visit(A, VOX(cp1, cq2), (rp - rq) * wbase);
            if (cp2 < np && cq2 < nq) ref_visit(A, VOX(cp2, cq2), (1 - rp) * wbase);
        } else {
            if (cp1 >= 0 && cq1 >= 0) ref_visit(A, VOX(cp1, cq1), rp * wbase);
            if (cp2 < np && cq1 >= 0) ref_visit(A, VOX(cp2, cq1), (rq - rp) * wbase);
            if (cp2 < np && cq2 < nq) ref_visit(A, VOX(cp2, cq2), (1 - rq) * wbase);
        }
    } else {
        /* corrected: walk from the p-low face; there q starts in cq2 and
         * crosses down into cq1 at fraction (1 - rq). */
        const float fq = 1 - rq;
        if (rp < fq) {
            if (cp1 >= 0 && cq2 < nq) ref_visit(A, VOX(cp1, cq2), rp * wbase);
            if (cp2 < np && cq2 < nq) ref_visit(A, VOX(cp2, cq2), (fq - rp) * wbase);
            if (cp2 < np && cq1 >= 0) ref_visit(A, VOX(cp2, cq1), (1 - fq) * wbase);
        } else {
            if (cp1 >= 0 && cq2 < nq) ref_visit(A, VOX(cp1, cq2), fq * wbase);
            if (cp1 >= 0 && cq1 >= 0) ref_visit(A, VOX(cp1, cq1), (rp - fq) * wbase);
            if (cp2 < np && cq1 >= 0) ref_visit(A, VOX(cp2, cq1), (1 - rp) * wbase);
        }
    }
#undef VOX
}

/* One ray of the reference walk (project_device / backproject_device). */
static void ref_ray(const vol32 *V, const float S[3], const float D[3], unsigned flags, int adjoint, ref_acc *A)
{
    const int nx = V->nx, ny = V->ny, nz = V->nz;
    const float dx_ = V->sx / nx;
    const float dy = V->sy / ny / dx_;
    const float dz = V->sz / nz / dx_;
    const float x1 = (S[0] - V->cx) / dx_, x2 = (D[0] - V->cx) / dx_;
    const float y1 = (S[1] - V->cy) / dx_, y2 = (D[1] - V->cy) / dx_;
    const float z1 = (S[2] - V->cz) / dx_, z2 = (D[2] - V->cz) / dx_;
    const float xd = x2 - x1, yd = y2 - y1, zd = z2 - z1;
    if (adjoint && (flags & ORC_SHORT_SKIP) && sqrtf(xd * xd + yd * yd) < 10.0f) return;
    const float nx2 = nx / 2.0f, ny2 = ny / 2.0f, nz2 = nz / 2.0f;
    const float L = sqrtf(xd * xd + yd * yd + zd * zd);
    const int refrule = (flags & ORC_REF_SLAB) != 0;
    const int xdom = (adjoint && (flags & ORC_TIE_GT)) ? (fabsf(xd) > fabsf(yd)) : (fabsf(xd) >= fabsf(yd));
    if (xdom) {
        const float ky = yd / xd, kz = zd / xd;
        float wbase = sqrtf(1 + ky * ky + kz * kz) * dx_;
        if (flags & ORC_INV_L2) wbase = wbase / L / L;
        const int flip = (ky >= 0.0f) != (kz >= 0.0f);
        for (int ix = 0; ix < nx; ++ix) {
            const float xx1 = ix - nx2, xx2 = xx1 + 1.0f;
            float yy1, yy2, zz1, zz2;
            if (ky >= 0.0f) { yy1 = (y1 + ky * (xx1 - x1)) / dy + ny2; yy2 = (y1 + ky * (xx2 - x1)) / dy + ny2; }
            else { yy1 = (y1 + ky * (xx2 - x1)) / dy + ny2; yy2 = (y1 + ky * (xx1 - x1)) / dy + ny2; }
            if (kz >= 0.0f) { zz1 = (z1 + kz * (xx1 - x1)) / dz + nz2; zz2 = (z1 + kz * (xx2 - x1)) / dz + nz2; }
            else { zz1 = (z1 + kz * (xx2 - x1)) / dz + nz2; zz2 = (z1 + kz * (xx1 - x1)) / dz + nz2; }
            ref_slab(A, yy1, yy2, zz1, zz2, ny, nz, ix, nx, (long)nx * ny, wbase, refrule, flip);
        }
    } else {
        const float kx = xd / yd, kz = zd / yd;
        float wbase = sqrtf(1 + kx * kx + kz * kz) * dx_;
        if (flags & ORC_INV_L2) wbase = wbase / L / L;
        const int flip = (kx >= 0.0f) != (kz >= 0.0f);
        for (int iy = 0; iy < ny; ++iy) {
            const float yy1 = iy - ny2, yy2 = yy1 + 1.0f;
            float xx1, xx2, zz1, zz2;
            if (kx >= 0.0f) { xx1 = (x1 + kx * (yy1 - y1)) + nx2; xx2 = (x1 + kx * (yy2 - y1)) + nx2; }
            else { xx1 = (x1 + kx * (yy2 - y1)) + nx2; xx2 = (x1 + kx * (yy1 - y1)) + nx2; }
            if (kz >= 0.0f) { zz1 = (z1 + kz * (yy1 - y1)) / dz + nz2; zz2 = (z1 + kz * (yy2 - y1)) / dz + nz2; }
            else { zz1 = (z1 + kz * (yy2 - y1)) / dz + nz2; zz2 = (z1 + kz * (yy1 - y1)) / dz + nz2; }
            ref_slab(A, xx1, xx2, zz1, zz2, nx, nz, (long)iy * nx, 1, (long)nx * ny, wbase, refrule, flip);
        }
    }
}

static void to_vol32(const orc_vol *V, vol32 *W)
{
    W->nx = V->nx; W->ny = V->ny; W->nz = V->nz;
    W->cx = (float)V->cx; W->cy = (float)V->cy; W->cz = (float)V->cz;
    W->sx = (float)V->sx; W->sy = (float)V->sy; W->sz = (float)V->sz;
}

/* Reference forward over views [v0, v0 + step*k) (all views: v0=0, step=1). */
void orc_ref_fwd(const orc_vol *V, const orc_det *G, const float *angles, int nv,
                 const float *img, float *out, unsigned flags, int v0, int vstep)
{
    vol32 W;
    to_vol32(V, &W);
    if (vstep <= 0) vstep = 1;
    const int nsel = (nv - v0 + vstep - 1) / vstep;
    const long per = (long)G->na * G->nb;
    const long nrays = per * nsel;
#pragma omp parallel for schedule(dynamic, 256)
    for (long r = 0; r < nrays; ++r) {
        const int ia = (int)(r % G->na), ib = (int)((r / G->na) % G->nb);
        const int iv = v0 + vstep * (int)(r / per);
        float S[3], D[3];
        ref_points32(G, angles[iv], ia, ib, S, D);
        ref_acc A = {img, NULL, 0.f, 0.f, 1};
        ref_ray(&W, S, D, flags, 0, &A);
        out[ia + (long)G->na * (ib + (long)G->nb * iv)] = A.acc;
    }
}

/* Reference adjoint; out is zeroed here (siddon.cc:265).  Threads own
 * private volumes over disjoint view subsets, then reduce. */
void orc_ref_bp(const orc_vol *V, const orc_det *G, const float *angles, int nv,
                const float *pv, float *out, unsigned flags, int v0, int vstep)
{
    vol32 W;
    to_vol32(V, &W);
    if (vstep <= 0) vstep = 1;
    const long nvox = (long)V->nx * V->ny * V->nz;
    memset(out, 0, sizeof(float) * nvox);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    float *priv = (float *)calloc((size_t)nthr * nvox, sizeof(float));
    if (!priv) return;
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        float *mine = priv + (size_t)t * nvox;
        const int nsel = (nv - v0 + vstep - 1) / vstep;
        const long rows = (long)nsel * G->nb;
#pragma omp for schedule(dynamic, 4)
        for (long row = 0; row < rows; ++row) {
            const int iv = v0 + vstep * (int)(row / G->nb), ib = (int)(row % G->nb);
                for (int ia = 0; ia < G->na; ++ia) {
                    float S[3], D[3];
                    ref_points32(G, angles[iv], ia, ib, S, D);
                    ref_acc A = {NULL, mine, pv[ia + (long)G->na * (ib + (long)G->nb * iv)], 0.f, 0};
                    ref_ray(&W, S, D, flags, 1, &A);
                }
        }
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nvox; ++i) {
        float acc = 0.0f;
        for (int t = 0; t < nthr; ++t) acc += priv[(size_t)t * nvox + i];
        out[i] = acc;
    }
    free(priv);
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
