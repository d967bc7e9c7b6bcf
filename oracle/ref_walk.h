/*
 * ref_walk.h -- restatement of the reference slab walk
 * (nefct/cc/siddon_3d/siddon.cu.cc:12-321 forward, :450-764 adjoint, ray set-up
 * :341-353 / :376-387), written once and included by oracle.c for REAL = float
 * (the reference's own precision) and REAL = double (the same rule without
 * FP32 rounding).  Test infrastructure only.
 */
/* ------------------------------------------------------------------ */
/* (2) FP32 restatement of the reference slab walk                     */
/* ------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz;
    REAL cx, cy, cz, sx, sy, sz;
} SFX(vol);

/* Reference ray set-up in REAL (siddon.cu.cc:341-353 / :376-387). */
static void SFX(ref_points)(const orc_det *g, REAL angle, int ia, int ib, REAL S[3], REAL D[3])
{
    const REAL na = (REAL)g->na, nb = (REAL)g->nb;
    const REAL da = (REAL)g->da, db = (REAL)g->db, ai = (REAL)g->ai, bi = (REAL)g->bi;
    const REAL SID = (REAL)g->sid, SAD = (REAL)g->sad;
    const REAL a = (ai + ia) * da - na * da / 2;
    const REAL b = (bi + ib) * db - nb * db / 2;
    const REAL c = COS(angle), s = SIN(angle);
    REAL lx, ly;
    if (g->cyl) {
        lx = SAD * COS(a) - SID;
        ly = SAD * SIN(a);
    } else {
        lx = SAD - SID;
        ly = a;
    }
    S[0] = -SID * c;
    S[1] = -SID * s;
    S[2] = ((REAL)0);
    D[0] = lx * c - ly * s;
    D[1] = lx * s + ly * c;
    D[2] = b;
}

/* Visitor for one weighted voxel: forward accumulates, adjoint scatters. */
typedef struct {
    const REAL *img; /* forward */
    REAL *out;       /* adjoint */
    REAL val;        /* adjoint ray value */
    REAL acc;        /* forward accumulator */
    int fwd;
} SFX(ref_acc);

static inline void SFX(ref_visit)(SFX(ref_acc) *A, long idx, REAL w)
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
static void SFX(ref_slab)(SFX(ref_acc) *A, REAL p1, REAL p2, REAL q1, REAL q2, int np, int nq,
                     long base, long sp, long sq, REAL wbase, int ref_slab_rule, int flip_q)
{
    const int cp1 = (int)FLOOR(p1), cp2 = (int)FLOOR(p2);
    const int cq1 = (int)FLOOR(q1), cq2 = (int)FLOOR(q2);
#define VOX(ip, iq) (base + (long)(ip) * sp + (long)(iq) * sq)
    if (cp1 == cp2) {
        if (!(0 <= cp1 && cp1 < np)) return;
        if (cq1 == cq2) {
            if (0 <= cq1 && cq1 < nq) SFX(ref_visit)(A, VOX(cp1, cq1), wbase);
        } else if (-1 <= cq1 && cq1 < nq) {
            const REAL rq = (cq2 - q1) / (q2 - q1);
            if (cq1 >= 0) SFX(ref_visit)(A, VOX(cp1, cq1), rq * wbase);
            if (cq2 < nq) SFX(ref_visit)(A, VOX(cp1, cq2), (1 - rq) * wbase);
        }
        return;
    }
    if (!(-1 <= cp1 && cp1 < np)) return;
    if (cq1 == cq2) {
        if (0 <= cq1 && cq1 < nq) {
            const REAL rp = (cp2 - p1) / (p2 - p1);
            if (cp1 >= 0) SFX(ref_visit)(A, VOX(cp1, cq1), rp * wbase);
            if (cp2 < np) SFX(ref_visit)(A, VOX(cp2, cq1), (1 - rp) * wbase);
        }
        return;
    }
    if (!(-1 <= cq1 && cq1 < nq)) return;
    const REAL rp = (cp2 - p1) / (p2 - p1);
    const REAL rq = (cq2 - q1) / (q2 - q1);
    if (ref_slab_rule || !flip_q) {
        /* reference attribution: both fractions from the low end (siddon.cu.cc:132-176) */
        if (rp > rq) {
            if (cp1 >= 0 && cq1 >= 0) SFX(ref_visit)(A, VOX(cp1, cq1), rq * wbase);
            if (cp1 >= 0 && cq2 < nq) SFX(ref_visit)(A, VOX(cp1, cq2), (rp - rq) * wbase);
            if (cp2 < np && cq2 < nq) SFX(ref_visit)(A, VOX(cp2, cq2), (1 - rp) * wbase);
        } else {
            if (cp1 >= 0 && cq1 >= 0) SFX(ref_visit)(A, VOX(cp1, cq1), rp * wbase);
            if (cp2 < np && cq1 >= 0) SFX(ref_visit)(A, VOX(cp2, cq1), (rq - rp) * wbase);
            if (cp2 < np && cq2 < nq) SFX(ref_visit)(A, VOX(cp2, cq2), (1 - rq) * wbase);
        }
    } else {
        /* corrected: walk from the p-low face; there q starts in cq2 and
         * crosses down into cq1 at fraction (1 - rq). */
        const REAL fq = 1 - rq;
        if (rp < fq) {
            if (cp1 >= 0 && cq2 < nq) SFX(ref_visit)(A, VOX(cp1, cq2), rp * wbase);
            if (cp2 < np && cq2 < nq) SFX(ref_visit)(A, VOX(cp2, cq2), (fq - rp) * wbase);
            if (cp2 < np && cq1 >= 0) SFX(ref_visit)(A, VOX(cp2, cq1), (1 - fq) * wbase);
        } else {
            if (cp1 >= 0 && cq2 < nq) SFX(ref_visit)(A, VOX(cp1, cq2), fq * wbase);
            if (cp1 >= 0 && cq1 >= 0) SFX(ref_visit)(A, VOX(cp1, cq1), (rp - fq) * wbase);
            if (cp2 < np && cq1 >= 0) SFX(ref_visit)(A, VOX(cp2, cq1), (1 - rp) * wbase);
        }
    }
#undef VOX
}

/* One ray of the reference walk (project_device / backproject_device). */
static void SFX(ref_ray)(const SFX(vol) *V, const REAL S[3], const REAL D[3], unsigned flags, int adjoint, SFX(ref_acc) *A)
{
    const int nx = V->nx, ny = V->ny, nz = V->nz;
    const REAL dx_ = V->sx / nx;
    const REAL dy = V->sy / ny / dx_;
    const REAL dz = V->sz / nz / dx_;
    const REAL x1 = (S[0] - V->cx) / dx_, x2 = (D[0] - V->cx) / dx_;
    const REAL y1 = (S[1] - V->cy) / dx_, y2 = (D[1] - V->cy) / dx_;
    const REAL z1 = (S[2] - V->cz) / dx_, z2 = (D[2] - V->cz) / dx_;
    const REAL xd = x2 - x1, yd = y2 - y1, zd = z2 - z1;
    if (adjoint && (flags & ORC_SHORT_SKIP) && SQRT(xd * xd + yd * yd) < ((REAL)10)) return;
    const REAL nx2 = nx / ((REAL)2), ny2 = ny / ((REAL)2), nz2 = nz / ((REAL)2);
    const REAL L = SQRT(xd * xd + yd * yd + zd * zd);
    const int refrule = (flags & ORC_REF_SLAB) != 0;
    const int xdom = (adjoint && (flags & ORC_TIE_GT)) ? (FABS(xd) > FABS(yd)) : (FABS(xd) >= FABS(yd));
    if (xdom) {
        const REAL ky = yd / xd, kz = zd / xd;
        REAL wbase = SQRT(1 + ky * ky + kz * kz) * dx_;
        if (flags & ORC_INV_L2) wbase = wbase / L / L;
        const int flip = (ky >= ((REAL)0)) != (kz >= ((REAL)0));
        for (int ix = 0; ix < nx; ++ix) {
            const REAL xx1 = ix - nx2, xx2 = xx1 + ((REAL)1);
            REAL yy1, yy2, zz1, zz2;
            if (ky >= ((REAL)0)) { yy1 = (y1 + ky * (xx1 - x1)) / dy + ny2; yy2 = (y1 + ky * (xx2 - x1)) / dy + ny2; }
            else { yy1 = (y1 + ky * (xx2 - x1)) / dy + ny2; yy2 = (y1 + ky * (xx1 - x1)) / dy + ny2; }
            if (kz >= ((REAL)0)) { zz1 = (z1 + kz * (xx1 - x1)) / dz + nz2; zz2 = (z1 + kz * (xx2 - x1)) / dz + nz2; }
            else { zz1 = (z1 + kz * (xx2 - x1)) / dz + nz2; zz2 = (z1 + kz * (xx1 - x1)) / dz + nz2; }
            SFX(ref_slab)(A, yy1, yy2, zz1, zz2, ny, nz, ix, nx, (long)nx * ny, wbase, refrule, flip);
        }
    } else {
        const REAL kx = xd / yd, kz = zd / yd;
        REAL wbase = SQRT(1 + kx * kx + kz * kz) * dx_;
        if (flags & ORC_INV_L2) wbase = wbase / L / L;
        const int flip = (kx >= ((REAL)0)) != (kz >= ((REAL)0));
        for (int iy = 0; iy < ny; ++iy) {
            const REAL yy1 = iy - ny2, yy2 = yy1 + ((REAL)1);
            REAL xx1, xx2, zz1, zz2;
            if (kx >= ((REAL)0)) { xx1 = (x1 + kx * (yy1 - y1)) + nx2; xx2 = (x1 + kx * (yy2 - y1)) + nx2; }
            else { xx1 = (x1 + kx * (yy2 - y1)) + nx2; xx2 = (x1 + kx * (yy1 - y1)) + nx2; }
            if (kz >= ((REAL)0)) { zz1 = (z1 + kz * (yy1 - y1)) / dz + nz2; zz2 = (z1 + kz * (yy2 - y1)) / dz + nz2; }
            else { zz1 = (z1 + kz * (yy2 - y1)) / dz + nz2; zz2 = (z1 + kz * (yy1 - y1)) / dz + nz2; }
            SFX(ref_slab)(A, xx1, xx2, zz1, zz2, nx, nz, (long)iy * nx, 1, (long)nx * ny, wbase, refrule, flip);
        }
    }
}

static void SFX(to_vol)(const orc_vol *V, SFX(vol) *W)
{
    W->nx = V->nx; W->ny = V->ny; W->nz = V->nz;
    W->cx = (REAL)V->cx; W->cy = (REAL)V->cy; W->cz = (REAL)V->cz;
    W->sx = (REAL)V->sx; W->sy = (REAL)V->sy; W->sz = (REAL)V->sz;
}

/* Reference forward over views [v0, v0 + step*k) (all views: v0=0, step=1). */
void SFX(orc_ref_fwd)(const orc_vol *V, const orc_det *G, const REAL *angles, int nv,
                 const REAL *img, REAL *out, unsigned flags, int v0, int vstep)
{
    SFX(vol) W;
    SFX(to_vol)(V, &W);
    if (vstep <= 0) vstep = 1;
    const int nsel = (nv - v0 + vstep - 1) / vstep;
    const long per = (long)G->na * G->nb;
    const long nrays = per * nsel;
#pragma omp parallel for schedule(dynamic, 256)
    for (long r = 0; r < nrays; ++r) {
        const int ia = (int)(r % G->na), ib = (int)((r / G->na) % G->nb);
        const int iv = v0 + vstep * (int)(r / per);
        REAL S[3], D[3];
        SFX(ref_points)(G, angles[iv], ia, ib, S, D);
        SFX(ref_acc) A = {img, NULL, 0.f, 0.f, 1};
        SFX(ref_ray)(&W, S, D, flags, 0, &A);
        out[ia + (long)G->na * (ib + (long)G->nb * iv)] = A.acc;
    }
}

/* Reference adjoint; out is zeroed here (siddon.cc:265).  Threads own
 * private volumes over disjoint view subsets, then reduce. */
void SFX(orc_ref_bp)(const orc_vol *V, const orc_det *G, const REAL *angles, int nv,
                const REAL *pv, REAL *out, unsigned flags, int v0, int vstep)
{
    SFX(vol) W;
    SFX(to_vol)(V, &W);
    if (vstep <= 0) vstep = 1;
    const long nvox = (long)V->nx * V->ny * V->nz;
    memset(out, 0, sizeof(REAL) * nvox);
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    REAL *priv = (REAL *)calloc((size_t)nthr * nvox, sizeof(REAL));
    if (!priv) return;
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        REAL *mine = priv + (size_t)t * nvox;
        const int nsel = (nv - v0 + vstep - 1) / vstep;
        const long rows = (long)nsel * G->nb;
#pragma omp for schedule(dynamic, 4)
        for (long row = 0; row < rows; ++row) {
            const int iv = v0 + vstep * (int)(row / G->nb), ib = (int)(row % G->nb);
                for (int ia = 0; ia < G->na; ++ia) {
                    REAL S[3], D[3];
                    SFX(ref_points)(G, angles[iv], ia, ib, S, D);
                    SFX(ref_acc) A = {NULL, mine, pv[ia + (long)G->na * (ib + (long)G->nb * iv)], 0.f, 0};
                    SFX(ref_ray)(&W, S, D, flags, 1, &A);
                }
        }
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nvox; ++i) {
        REAL acc = ((REAL)0);
        for (int t = 0; t < nthr; ++t) acc += priv[(size_t)t * nvox + i];
        out[i] = acc;
    }
    free(priv);
}

