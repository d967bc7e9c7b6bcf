// Cone-beam Siddon projector pair for gfx950.
//
// Replaces nefct/cc/siddon_3d/siddon.cu.cc (project_device :12-321,
// ProjectFlat/CyliKernel :323-390, backproject_device :450-764,
// BackProjectFlat/CyliKernel :766-832, launchers :392-448, :834-890).
//
// Design (one lane = one ray, ray-driven, slab-synchronous):
//  * per-ray set-up in FP64: end points, dominant axis, transverse slopes,
//    entry/exit slab range clipped to the volume box, mm length per slab;
//  * the slab loop runs in FP32 over wave-uniform slab indices (min/max over
//    the wave's rays), so at every iteration the 64 lanes of a wavefront sit
//    in the same slab;
//  * the volume is re-laid out per call so that the transverse axis that
//    varies across detector columns is the fastest one: x-slab walks read a
//    y-fastest copy, y-slab walks an x-fastest copy.  Both copies carry a
//    2-voxel zero border on their transverse axes so the slab kernel needs
//    no per-voxel bounds test;
//  * every volume access goes through a raw buffer resource sized to the
//    copy, so an index error reads 0 / drops the atomic instead of faulting;
//  * the adjoint scatters with buffer float atomics into the same two
//    layouts, then one pass sums them into the x-fastest output.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <initializer_list>
#include <mutex>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "common.h"

namespace {

constexpr int PAD = 2;          // zero border of the transverse axes
constexpr int BX = 64;          // window adjoint: detector columns per wavefront
constexpr int BY = 4;           // window adjoint: detector rows per workgroup (one per wave)
constexpr int FX = 16;          // forward: detector columns per workgroup (x fastest: a wavefront is 16 x 4 rays)
constexpr int FY = 16;          // forward: detector rows per workgroup

struct VolDev {
    int n[3];
    double lo[3];
    double vs[3];
    double inv_vs[3];
};

struct DetDev {
    int cyl;
    double sid, sad;
    int na, nb;
    double da, ai, db, bi;
};

// Element offsets of a padded layout for one walk direction.
//   elem(ca, cb, t) = ca*sa + cb*sb + t*st + org   (ca, cb centred cells)
struct Walk {
    int sa, sb, st, org;
    int na, nb;      // transverse extents (for checked walks / counting)
    int ha, hb;      // centring offsets: true cell = centred + h
};

struct Layouts {
    float* xf;       // (nx+2P) x ny x (nz+2P), x fastest
    float* yf;       // (ny+2P) x nx x (nz+2P), y fastest
    uint32_t xf_bytes, yf_bytes;
    int NXP, NYP, NZP;
};

struct Ray {
    int dom;             // 0: x slabs, 1: y slabs, 2: z slabs, -1: misses
    int s0, s1;          // slab range [s0, s1)
    int tc;              // centre slab
    float ka, kb;        // transverse slopes (cells per slab)
    float Ac, Bc;        // centred transverse coordinates at slab face tc
    float inv_ka, inv_kb;
    float pa, pb;        // 1 if slope > 0 else 0
    float scale;         // mm per unit slab step (x 1/L^2 with NEFCT_F_INV_L2)
};

__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// Cell of a walk coordinate and the position inside it: floor(x) as an int
// and x - floor(x) (v_fract_f32).  The coordinates are bounded (|x| < 2^20).
// This file's device code is compiled without NaN semantics
// (-Xarch_device -fno-honor-nans), which lets (int)floorf(x) become one
// v_cvt_flr_i32_f32 (~5 % of the walk loops).  Nothing here relies on NaN
// arithmetic: tests for non-finite data work on the bits (fixed_exp,
// the fixed-point write-out, the window adjoint's flush), and walk and
// set-up values are finite by construction.
__device__ __forceinline__ int cell_floor(float x) { return (int)floorf(x); }
__device__ __forceinline__ float cell_frac(float x) { return __builtin_amdgcn_fractf(x); }

// Exact rule, one transverse axis: fraction of the slab walked before the
// coordinate leaves its cell, (pa - frac) / k = fma(frac, -1/k, pa/k) with
// pa = 1 for k > 0 (next boundary above) and 0 for k < 0, clamped to [0, 1].
// It is >= 1 when the coordinate stays in its cell, so the clamp alone gives
// 1 there; rounding can leave it just below 1 without a crossing, which
// splits the slab's length between two pieces of the same cell (sum
// unchanged).
__device__ __forceinline__ float cross_frac(float frac, float inv_k, float pa_inv_k)
{
    return clamp01(fmaf(frac, -inv_k, pa_inv_k));
}

// Rays with a transverse slope within 2^-8 of +-1 (views near 45 degrees).
// Their FP32 face coordinates advance by up to |k| + 2 ulp per slab, so a
// face coordinate can round across a second cell boundary (with |x| < 2^13
// one ulp is <= 2^-10) and the cells taken from floor() would skip a cell.
// These rays take the cell sequence from the crossings instead: one step per
// crossing, the in-cell fraction measured from the stepped cell (it may leave
// [0, 1] by a rounding error, which the clamp of cross_frac absorbs: a
// coordinate past the boundary crosses at once, one short of it does not).
constexpr float K_NEAR_ONE = 1.0f - 1.0f / 256.0f;
__device__ __forceinline__ bool near_unit_slope(const Ray& r)
{
    return fabsf(r.ka) > K_NEAR_ONE || fabsf(r.kb) > K_NEAR_ONE;
}

// floor / ceil of v as an int clamped to [lo, hi] before the conversion, so
// huge values (slopes near zero) cannot overflow; v is finite
__device__ __forceinline__ int floor_clamp(float v, int lo, int hi)
{
    return (int)fminf(fmaxf(floorf(v), (float)lo), (float)hi);
}
__device__ __forceinline__ int ceil_clamp(float v, int lo, int hi)
{
    return (int)fminf(fmaxf(ceilf(v), (float)lo), (float)hi);
}

// The ray set-up (ray_points, clip_axis, setup_ray_g) is evaluated by the
// forward, the adjoints and the counter in different inlining contexts; with
// FMA contraction off and the fused operations written out, every context
// rounds identically, which keeps the pair an exact transpose.
__device__ __forceinline__ void ray_points(const DetDev& g, double c, double s, int ia, int ib,
                                           double S[3], double D[3], double dz = 0.0)
{
#pragma clang fp contract(off)
    const double a = fma(g.ai + ia, g.da, -(g.na * g.da * 0.5));
    const double b = fma(g.bi + ib, g.db, -(g.nb * g.db * 0.5));
    double lx, ly;
    if (g.cyl) {
        lx = g.sad * cos(a) - g.sid;
        ly = g.sad * sin(a);
    } else {
        lx = g.sad - g.sid;
        ly = a;
    }
    S[0] = -g.sid * c;
    S[1] = -g.sid * s;
    S[2] = dz;
    D[0] = fma(lx, c, -(ly * s));
    D[1] = fma(lx, s, ly * c);
    D[2] = b + dz;
}

// Restricts [tin, tout] to the slab parameters where lo <= A + k t <= hi
// (inv_k = 1/k, unused when k == 0).
__device__ __forceinline__ void clip_axis(double k, double inv_k, double A, double lo, double hi, double& tin,
                                          double& tout)
{
#pragma clang fp contract(off)
    if (k == 0.0) {
        if (!(A > lo && A < hi)) { tin = 1.0; tout = 0.0; }
        return;
    }
    double t0 = (lo - A) * inv_k, t1 = (hi - A) * inv_k;
    if (t0 > t1) { double t = t0; t0 = t1; t1 = t; }
    tin = fmax(tin, t0);
    tout = fmin(tout, t1);
}


// What the set-up needs of walk axis m and its transverse axes a, b: the
// volume's extents and voxel sizes and the view's source in voxel units.
struct AxisGeom {
    double vsm, vsa, vsb, ivsm;
    double uSm, uSa, uSb;
    int nm, na, nb;
    int pad_[17];   // 128 bytes: a table entry's LDS offset is a shift of the axis index
};

// The view's source in voxel units (the S of ray_points).
__device__ __forceinline__ void source_voxels(const VolDev& V, const DetDev& G, double c, double s, double uS[3],
                                              double dz = 0.0)
{
    const double S[3] = {-G.sid * c, -G.sid * s, dz};
    for (int i = 0; i < 3; ++i) uS[i] = (S[i] - V.lo[i]) * V.inv_vs[i];
}

// FP64 ray set-up (three reciprocals; the length per slab in FP32).
// 'ref' selects the reference's x/y-only dominance rule.  geom(m) supplies the
// AxisGeom of the ray's walk axis (selected from V per lane, or read from a
// table), with the same values either way.
// EARLY_EXIT returns at a miss with the fields zeroed (fewer registers: the
// forward keeps 8 waves per SIMD); without it the set-up is straight-line
// (fewer instructions in the adjoint's per-block set-up).  Both give the same
// fields for a ray that hits.
template <bool EARLY_EXIT = false, class Geom>
__device__ __forceinline__ void setup_ray_g(const VolDev& V, const DetDev& G, double c, double s, int ia, int ib,
                                            unsigned flags, bool adjoint, Geom&& geom, Ray& r, double dz = 0.0)
{
#pragma clang fp contract(off)
    double S[3], D[3], d[3];
    ray_points(G, c, s, ia, ib, S, D, dz);
    for (int i = 0; i < 3; ++i) d[i] = (D[i] - S[i]) * V.inv_vs[i];
    if constexpr (EARLY_EXIT) {
        r.dom = -1;
        r.s0 = r.s1 = 0;
        r.tc = 0;
        r.ka = r.kb = r.Ac = r.Bc = r.inv_ka = r.inv_kb = r.pa = r.pb = 0.0f;
        r.scale = 0.0f;
    }
    int m;
    bool hit = true;
    const bool ref = (flags & NEFCT_F_REF_SLAB) != 0;
    // reference quantities in x-voxel units (siddon.cu.cc:19-38)
    const double xd = (D[0] - S[0]) * V.inv_vs[0], yd = (D[1] - S[1]) * V.inv_vs[0], zd = (D[2] - S[2]) * V.inv_vs[0];
    if (ref) {
        if (adjoint && (flags & NEFCT_F_SHORT_SKIP) && sqrt(fma(yd, yd, xd * xd)) < 10.0) hit = false;
        if (EARLY_EXIT && !hit) return;
        const bool xdom = (adjoint && (flags & NEFCT_F_TIE_GT)) ? (fabs(xd) > fabs(yd)) : (fabs(xd) >= fabs(yd));
        m = xdom ? 0 : 1;
    } else {
        const double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
        m = (ax >= ay && ax >= az) ? 0 : ((ay >= az) ? 1 : 2);
    }
    // Straight-line from here (without EARLY_EXIT): a ray that misses
    // (d[m] == 0, or no slab of the box) computes on with values that are not
    // used, and the fields are written once at the end (early returns make the
    // compiler materialise every field's value on every exit path, a dozen
    // moves per ray).
    hit = hit && d[m] != 0.0;
    if (EARLY_EXIT && !hit) return;
    const int a = (m == 0) ? 1 : 0;
    const int b = (m == 2) ? 1 : 2;
    const AxisGeom g = geom(m);
    const double inv_dm = rcp64(d[m]);
    const double ka = d[a] * inv_dm, kb = d[b] * inv_dm;
    const double ika = (d[a] != 0.0) ? d[m] * rcp64(d[a]) : 0.0;
    const double ikb = (d[b] != 0.0) ? d[m] * rcp64(d[b]) : 0.0;
    const double Aa = fma(-ka, g.uSm, g.uSa), Ab = fma(-kb, g.uSm, g.uSb);
    // The exact walk clips to the volume box.  The reference rule can credit an
    // in-range voxel from a slab whose exact segment lies just outside the box
    // (the Q2 attribution at a corner), so reference mode clips to the box
    // widened by one cell on each transverse side; its per-voxel range tests
    // then decide exactly as siddon.cu.cc:80-176 does.  (tin and tout stay
    // finite: fmax and fmin drop NaN operands, and start from 0 and nm.)
    const double mg = ref ? 1.0 : 0.0;
    double tin = 0.0, tout = (double)g.nm;
    clip_axis(ka, ika, Aa, -mg, g.na + mg, tin, tout);
    clip_axis(kb, ikb, Ab, -mg, g.nb + mg, tin, tout);
    hit = hit && tout > tin;
    if (EARLY_EXIT && !hit) return;
    int s0 = (int)floor(tin), s1 = (int)ceil(tout);
    s0 = max(s0, 0);
    s1 = min(s1, g.nm);
    hit = hit && s1 > s0;
    if (EARLY_EXIT && !hit) return;
    const int tc = (s0 + s1) >> 1;
    const int ha = g.na >> 1, hb = g.nb >> 1;
    r.dom = hit ? m : -1;
    r.s0 = s0;
    r.s1 = s1;
    r.tc = tc;
    r.ka = (float)ka;
    r.kb = (float)kb;
    r.Ac = (float)(fma(ka, (double)tc, Aa) - ha);
    r.Bc = (float)(fma(kb, (double)tc, Ab) - hb);
    r.inv_ka = (float)ika;
    r.inv_kb = (float)ikb;
    r.pa = ka > 0.0 ? 1.0f : 0.0f;
    r.pb = kb > 0.0 ? 1.0f : 0.0f;
    const double qa = (ka * g.vsa) * g.ivsm, qb = (kb * g.vsb) * g.ivsm;
    // the hardware square root (within 1 ulp; the argument is >= 1), the
    // product rounded to FP32 below
    double sc = g.vsm * (double)__builtin_amdgcn_sqrtf((float)fma(qb, qb, fma(qa, qa, 1.0)));
    if (flags & NEFCT_F_INV_L2) sc = sc / fma(zd, zd, fma(yd, yd, xd * xd));
    // a missing ray's scale is 0: the forward writes 0 * scale for it
    r.scale = hit ? (float)sc : 0.0f;
}

// AxisGeom of walk axis m from the volume record.
__device__ __forceinline__ AxisGeom axis_geom(const VolDev& V, const double uS[3], int m)
{
    const int a = (m == 0) ? 1 : 0, b = (m == 2) ? 1 : 2;
    AxisGeom g;
    g.nm = V.n[m];
    g.na = V.n[a];
    g.nb = V.n[b];
    g.vsm = V.vs[m];
    g.vsa = V.vs[a];
    g.vsb = V.vs[b];
    g.ivsm = V.inv_vs[m];
    g.uSm = uS[m];
    g.uSa = uS[a];
    g.uSb = uS[b];
    return g;
}

__device__ __forceinline__ void setup_ray(const VolDev& V, const DetDev& G, double c, double s, int ia, int ib,
                                          unsigned flags, bool adjoint, Ray& r, double dz = 0.0)
{
    double uS[3];
    source_voxels(V, G, c, s, uS, dz);
    setup_ray_g<true>(V, G, c, s, ia, ib, flags, adjoint, [&](int m) { return axis_geom(V, uS, m); }, r, dz);
}

__device__ __forceinline__ Walk walk_of(int dom, const VolDev& V, const Layouts& Ly)
{
    Walk w;
    const int nx = V.n[0], ny = V.n[1], nz = V.n[2];
    if (dom == 0) {            // x slabs: a = y (fast), b = z; y-fastest copy
        w.sa = 1;
        w.sb = Ly.NYP * nx;
        w.st = Ly.NYP;
        w.na = ny;
        w.nb = nz;
        w.ha = ny >> 1;
        w.hb = nz >> 1;
        w.org = (w.ha + PAD) + (w.hb + PAD) * w.sb;
    } else if (dom == 1) {     // y slabs: a = x (fast), b = z; x-fastest copy
        w.sa = 1;
        w.sb = Ly.NXP * ny;
        w.st = Ly.NXP;
        w.na = nx;
        w.nb = nz;
        w.ha = nx >> 1;
        w.hb = nz >> 1;
        w.org = (w.ha + PAD) + (w.hb + PAD) * w.sb;
    } else {                   // z slabs: a = x, b = y (unpadded: checked); x-fastest copy
        w.sa = 1;
        w.sb = Ly.NXP;
        w.st = Ly.NXP * ny;
        w.na = nx;
        w.nb = ny;
        w.ha = nx >> 1;
        w.hb = ny >> 1;
        w.org = (w.ha + PAD) + w.hb * w.sb + PAD * w.st;
    }
    return w;
}

// Per-slab segment decomposition.  Returns the three (centred) cells and
// their fractions of the slab.  The exact rule walks from the slab's lower
// face (slab_exact: entry cell, the cell after the first crossing, exit
// cell); the reference rule (NEFCT_F_REF_SLAB, slab_segments) measures each
// axis from its low-value end (siddon.cu.cc:52-76, 132-176).  GUARD tests the
// index change explicitly instead of relying on the clamp (cross_frac), for
// counting voxels (each once).
struct Seg {
    int a1, b1, a2, b2, a3, b3;
    float w1, w2, w3;
    bool ok;     // reference outer range test (siddon.cu.cc:80,112,132)
};

// the three pieces of a slab from its crossing fractions and cells
__device__ __forceinline__ void slab_pieces(int ca0, int ca1, float fa, int cb0, int cb1, float fb, Seg& s)
{
    const bool afirst = fa < fb;
    const float f1 = afirst ? fa : fb, f2 = afirst ? fb : fa;
    s.a1 = ca0; s.b1 = cb0;
    s.a2 = afirst ? ca1 : ca0;
    s.b2 = afirst ? cb0 : cb1;
    s.a3 = ca1; s.b3 = cb1;
    s.w1 = f1; s.w2 = f2 - f1; s.w3 = 1.0f - f2;
    s.ok = true;
}

template <bool GUARD>
__device__ __forceinline__ void slab_exact(int ca0, int ca1, float fra0, int cb0, int cb1, float frb0, const Ray& r,
                                           float pia, float pib, Seg& s)
{
    float fa = cross_frac(fra0, r.inv_ka, pia);
    float fb = cross_frac(frb0, r.inv_kb, pib);
    if (GUARD) {
        fa = (ca1 != ca0) ? fa : 1.0f;
        fb = (cb1 != cb0) ? fb : 1.0f;
    }
    slab_pieces(ca0, ca1, fa, cb0, cb1, fb, s);
}

// near_unit_slope rays: the cells step with the crossings; returns the exit
// cells and the in-cell fractions at the slab's far face
__device__ __forceinline__ void slab_step(int ca0, float fra0, int cb0, float frb0, float a1, float b1, const Ray& r,
                                          float pia, float pib, int& ca1, float& fra1, int& cb1, float& frb1, Seg& s)
{
    const float fa = (r.ka != 0.0f) ? cross_frac(fra0, r.inv_ka, pia) : 1.0f;
    const float fb = (r.kb != 0.0f) ? cross_frac(frb0, r.inv_kb, pib) : 1.0f;
    ca1 = ca0 + ((fa < 1.0f) ? ((r.ka > 0.0f) ? 1 : -1) : 0);
    cb1 = cb0 + ((fb < 1.0f) ? ((r.kb > 0.0f) ? 1 : -1) : 0);
    fra1 = a1 - (float)ca1;
    frb1 = b1 - (float)cb1;
    slab_pieces(ca0, ca1, fa, cb0, cb1, fb, s);
}

__device__ __forceinline__ void slab_segments(float a0, float a1, float fla0, float fla1, float b0, float b1,
                                              float flb0, float flb1, const Walk& w, Seg& s)
{
    const int ca0 = (int)fla0, ca1 = (int)fla1, cb0 = (int)flb0, cb1 = (int)flb1;
    const float loa = fminf(a0, a1), hia = fmaxf(a0, a1);
    const float lob = fminf(b0, b1), hib = fmaxf(b0, b1);
    const int cla = min(ca0, ca1), cha = max(ca0, ca1);
    const int clb = min(cb0, cb1), chb = max(cb0, cb1);
    const float ra = (cla != cha) ? ((float)cha - loa) / (hia - loa) : 1.0f;
    const float rb = (clb != chb) ? ((float)chb - lob) / (hib - lob) : 1.0f;
    const int tla = cla + w.ha, tlb = clb + w.hb;
    const bool oka = (cla == cha) ? (tla >= 0 && tla < w.na) : (tla >= -1 && tla < w.na);
    const bool okb = (clb == chb) ? (tlb >= 0 && tlb < w.nb) : (tlb >= -1 && tlb < w.nb);
    const float f1 = fminf(ra, rb), f2 = fmaxf(ra, rb);
    s.a1 = cla; s.b1 = clb;
    s.a2 = (ra > rb) ? cla : cha;
    s.b2 = (ra > rb) ? chb : clb;
    s.a3 = cha; s.b3 = chb;
    s.w1 = f1; s.w2 = f2 - f1; s.w3 = 1.0f - f2;
    s.ok = oka && okb;
}

__device__ __forceinline__ bool cell_in(const Walk& w, int ca, int cb)
{
    return (unsigned)(ca + w.ha) < (unsigned)w.na && (unsigned)(cb + w.hb) < (unsigned)w.nb;
}

// MODE 0: forward (returns the weighted sum), 1: adjoint scatter of 'val',
// 2: count of voxels with nonzero length.
//
// The three gathers of a slab are issued as three predicated loads whose
// results are consumed only after all of them, so a wavefront has all of a
// slab's loads (up to four slabs' with the unroll) in flight at once.
template <int MODE, bool REF, bool CHECKED, bool STEP = false>
__device__ float walk_ray(const Ray& r, bool mine, int tmin, int tmax, const Walk& w,
                          __amdgpu_buffer_rsrc_t rs, float val)
{
    float acc = 0.0f;
    float tf = (float)(tmin - r.tc);
    float a0 = fmaf(r.ka, tf, r.Ac), b0 = fmaf(r.kb, tf, r.Bc);
    const float sv = val * r.scale;
    int tb = tmin * w.st + w.org;
    // per slab: entry state (a0/b0 and their cells) carried from the previous slab
    float fla0 = floorf(a0), flb0 = floorf(b0);               // reference rule
    int ca0 = cell_floor(a0), cb0 = cell_floor(b0);           // exact rule
    float fra0 = cell_frac(a0), frb0 = cell_frac(b0);
    const float pia = r.pa * r.inv_ka, pib = r.pb * r.inv_kb;
    auto slab = [&](int t, auto ranged_tag) {
        constexpr bool RANGED = decltype(ranged_tag)::value;
        tf += 1.0f;
        const float a1 = fmaf(r.ka, tf, r.Ac), b1 = fmaf(r.kb, tf, r.Bc);
        Seg s;
        float fla1 = 0.0f, flb1 = 0.0f, fra1 = 0.0f, frb1 = 0.0f;
        int ca1 = 0, cb1 = 0;
        if (REF) {
            fla1 = floorf(a1);
            flb1 = floorf(b1);
            slab_segments(a0, a1, fla0, fla1, b0, b1, flb0, flb1, w, s);
        } else if (STEP) {
            slab_step(ca0, fra0, cb0, frb0, a1, b1, r, pia, pib, ca1, fra1, cb1, frb1, s);
        } else {
            ca1 = cell_floor(a1);
            cb1 = cell_floor(b1);
            fra1 = cell_frac(a1);
            frb1 = cell_frac(b1);
            slab_exact<MODE == 2>(ca0, ca1, fra0, cb0, cb1, frb0, r, pia, pib, s);
        }
        const bool act = mine && s.ok && (!RANGED || ((t >= r.s0) && (t < r.s1)));
        const int o1 = (s.a1 + __mul24(s.b1, w.sb) + tb) * 4;
        const int o2 = (s.a2 + __mul24(s.b2, w.sb) + tb) * 4;
        const int o3 = (s.a3 + __mul24(s.b3, w.sb) + tb) * 4;
        bool u1 = act && s.w1 > 0.0f, u2 = act && s.w2 > 0.0f, u3 = act && s.w3 > 0.0f;
        if (CHECKED || REF || MODE == 2) {
            u1 = u1 && cell_in(w, s.a1, s.b1);
            u2 = u2 && cell_in(w, s.a2, s.b2);
            u3 = u3 && cell_in(w, s.a3, s.b3);
        }
        if (MODE == 0) {
            float v1 = 0.0f, v2 = 0.0f, v3 = 0.0f;
            if (u1) v1 = buf_load(rs, o1);
            if (u2) v2 = buf_load(rs, o2);
            if (u3) v3 = buf_load(rs, o3);
            acc = fmaf(s.w1, v1, acc);
            acc = fmaf(s.w2, v2, acc);
            acc = fmaf(s.w3, v3, acc);
        } else if (MODE == 1) {
            if (u1) buf_atomic_add(rs, o1, s.w1 * sv);
            if (u2) buf_atomic_add(rs, o2, s.w2 * sv);
            if (u3) buf_atomic_add(rs, o3, s.w3 * sv);
        } else {
            acc += (float)((int)u1 + (int)u2 + (int)u3);
        }
        a0 = a1; b0 = b1;
        if (REF) {
            fla0 = fla1; flb0 = flb1;
        } else {
            ca0 = ca1; cb0 = cb1; fra0 = fra1; frb0 = frb1;
        }
        tb += w.st;
    };
    if (MODE == 0) {
        // the forward splits the wavefront's slab range: in [m0, m1) every
        // walking ray is inside its own range [s0, s1), so only the ends test it
        const int m0 = __builtin_amdgcn_readfirstlane(wave_max_i32(mine ? r.s0 : tmin));
        const int m1 = __builtin_amdgcn_readfirstlane(wave_min_i32(mine ? r.s1 : tmax));
        const int ta = min(max(m0, tmin), tmax), tb2 = max(min(m1, tmax), ta);
        // the middle part is unrolled by 4: four slabs' twelve gathers in
        // flight per trip (2: 5 % slower at 720 views, 8: 25 % slower)
#pragma unroll 2
        for (int t = tmin; t < ta; ++t) slab(t, std::true_type{});
#pragma unroll 4
        for (int t = ta; t < tb2; ++t) slab(t, std::false_type{});
#pragma unroll 2
        for (int t = tb2; t < tmax; ++t) slab(t, std::true_type{});
    } else {
#pragma unroll 2
        for (int t = tmin; t < tmax; ++t) slab(t, std::true_type{});
    }
    return acc;
}

struct Launch {
    VolDev V;
    DetDev G;
    Layouts Ly;
    const double* trig;  // (cos, sin) per view
    const float* zoff;   // per-view z offset of source and detector (mm), or null
    int nv;
    unsigned flags;
    float sbound;        // upper bound of any ray's mm per slab (fixed-point adjoint scale)
};

// A helical view's lift: source and detector of view iv raised by zoff[iv]
// along z (dist_3d_back_fdk.cu.cc:46,65); 0 without offsets.  (Passed to the
// set-up as a value: a per-view copy of the volume record would be indexed
// by axis at run time and so be placed in scratch memory.)
__device__ __forceinline__ double view_lift(const Launch& P, int iv)
{
    return P.zoff ? (double)P.zoff[iv] : 0.0;
}

// Upper bound of Ray::scale over every ray of the geometry (mm of ray per
// slab, times 1/L^2 under INV_L2), with 1 % margin.  Exact mode: the walk axis
// is the ray's largest component in voxel units, so the length per slab is at
// most the voxel diagonal.  Reference mode walks x or y: the in-plane part is
// at least 1/sqrt(2) of the in-plane length, and the source-detector vector
// rises at most |b|max over an in-plane length of at least SAD.
__host__ __device__ inline float scale_bound(const VolDev& V, const DetDev& G, unsigned flags)
{
    double bound;
    if (!(flags & NEFCT_F_REF_SLAB)) {
        bound = sqrt(V.vs[0] * V.vs[0] + V.vs[1] * V.vs[1] + V.vs[2] * V.vs[2]);
    } else {
        const double b_lo = G.bi * G.db - G.nb * G.db * 0.5, b_hi = (G.bi + G.nb - 1) * G.db - G.nb * G.db * 0.5;
        const double bmax = fmax(fabs(b_lo), fabs(b_hi)) + G.db;
        const double q = bmax / G.sad;
        bound = fmax(V.vs[0], V.vs[1]) * sqrt(2.0) * sqrt(1.0 + q * q);
    }
    if (flags & NEFCT_F_INV_L2) {
        const double lmin = G.sad * V.inv_vs[0];   // shortest source-detector distance, x-voxel units
        if (lmin < 1.0) bound /= lmin * lmin;
    }
    return (float)(bound * 1.01);
}

template <int MODE, bool REF>
__device__ float project_all_dominances(const Launch& P, const Ray& r, bool inside, float val)
{
    float acc = 0.0f;
    // x slabs and y slabs in the padded copies; z slabs checked in the x copy
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (REF && d == 2) break;
        const bool mine = inside && r.dom == d;
        const int lo = __builtin_amdgcn_readfirstlane(wave_min_i32(mine ? r.s0 : 0x7fffffff));
        const int hi = __builtin_amdgcn_readfirstlane(wave_max_i32(mine ? r.s1 : -0x7fffffff));
        if (lo >= hi) continue;
        const Walk w = walk_of(d, P.V, P.Ly);
        float* base = (d == 0) ? P.Ly.yf : P.Ly.xf;
        const uint32_t nbytes = (d == 0) ? P.Ly.yf_bytes : P.Ly.xf_bytes;
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(base, nbytes);
        // rays of near-unit slope walk the stepped-cell rule (a few per view
        // near 45 degrees); the branch costs nothing when no lane has one
        const bool stp = !REF && mine && near_unit_slope(r);
        float v = 0.0f;
        if (__any(mine && !stp)) {
            if (d == 2)
                v = walk_ray<MODE, REF, true>(r, mine && !stp, lo, hi, w, rs, val);
            else
                v = walk_ray<MODE, REF, false>(r, mine && !stp, lo, hi, w, rs, val);
        }
        if (!REF && __any(stp)) {
            float v2;
            if (d == 2)
                v2 = walk_ray<MODE, REF, true, true>(r, stp, lo, hi, w, rs, val);
            else
                v2 = walk_ray<MODE, REF, false, true>(r, stp, lo, hi, w, rs, val);
            if (stp) v = v2;
        }
        if (mine) acc = v;
    }
    return acc;
}

// One workgroup of FX x FY rays (columns ia, rows ib) of view iv.
template <int MODE, bool REF>
__device__ __forceinline__ void siddon_rays(const Launch& P, int ia, int ib, int iv, const float* __restrict__ pv_in,
                                            float* __restrict__ pv_out, int32_t* __restrict__ counts)
{
    const bool inside = ia < P.G.na && ib < P.G.nb;
    const double c = P.trig[2 * iv], s = P.trig[2 * iv + 1];
    Ray r;
    r.dom = -1;
    r.s0 = r.s1 = r.tc = 0;
    r.ka = r.kb = r.Ac = r.Bc = r.inv_ka = r.inv_kb = r.pa = r.pb = 0.0f;
    r.scale = 0.0f;
    if (inside) setup_ray(P.V, P.G, c, s, ia, ib, P.flags, MODE == 1, r, view_lift(P, iv));
    const size_t id = (size_t)ia + (size_t)P.G.na * ((size_t)ib + (size_t)P.G.nb * (size_t)iv);
    float val = 0.0f;
    if (MODE == 1 && inside) val = pv_in[id];
    const float acc = project_all_dominances<MODE, REF>(P, r, inside, val);
    if (!inside) return;
    if (MODE == 0) pv_out[id] = acc * r.scale;
    if (MODE == 2) counts[id] = (int32_t)acc;
}

// Workgroup order of the forward.  Consecutive workgroup ids go to the 8
// XCDs in turn (id mod 8), each XCD with its own L2.  The 1-D launch is
// remapped so that every XCD gets a contiguous range of (ray tile, view)
// tasks in which the view varies fastest within groups of 2^vg_log2 views:
// the rays an XCD has in flight then pass through nearly the same voxels (the
// same detector tile seen from neighbouring angles), and their gathers hit
// that XCD's L2 instead of re-reading the volume from HBM for every view.
// The group spans about 16 degrees of a full scan (fwd_group_log2).
constexpr int FWD_VG = 16;   // view group of the batched forward

__host__ inline int fwd_group_log2(int nv)
{
    int g = 2;                                  // 4 views at least ...
    while (g < 5 && (nv >> (g + 1)) >= 22) ++g; // ... up to 32 while 2^g views span <= 360/22 degrees
    return g;
}

// TX x TY rays per workgroup (256 threads, x fastest: a wavefront is TX x
// 64/TX rays); 16 x 16 by default, the others selectable by the forward tile
// flags for the autotuner.
template <int MODE, bool REF, int TX = FX, int TY = FY>
__global__ __launch_bounds__(TX * TY) void k_siddon(Launch P, const float* __restrict__ pv_in,
                                                     float* __restrict__ pv_out, int32_t* __restrict__ counts,
                                                     int tiles_a, int tiles_b, int v0, int nvc, int vg_log2)
{
    const unsigned per_xcd = gridDim.x >> 3;      // the grid is a multiple of 8
    const unsigned task = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);
    const unsigned tiles = (unsigned)tiles_a * (unsigned)tiles_b;
    const unsigned vin = task & ((1u << vg_log2) - 1u), rest = task >> vg_log2;
    const unsigned tile = rest % tiles, group = rest / tiles;
    const int iv = (int)((group << vg_log2) + vin);
    if (iv >= nvc) return;
    const int ta = (int)(tile % (unsigned)tiles_a), tb = (int)(tile / (unsigned)tiles_a);
    siddon_rays<MODE, REF>(P, ta * TX + threadIdx.x, tb * TY + threadIdx.y, v0 + iv, pv_in, pv_out, counts);
}

// View group of the forward's work order: log2 of the views an XCD walks
// together, from the flags (NEFCT_F_FWD_VG_*) or the default rule.
__host__ inline int fwd_vg_log2(unsigned flags, int nv)
{
    const unsigned v = (flags & NEFCT_F_FWD_VG_MASK) >> NEFCT_F_FWD_VG_SHIFT;
    return v ? (int)v : fwd_group_log2(nv);
}

// Launches k_siddon over all views in chunks whose 1-D grids stay below 2^31
// work-items.
template <int MODE, bool REF, int TX, int TY>
void launch_siddon_tiles(const Launch& L, const float* pv_in, float* pv_out, int32_t* counts, hipStream_t s)
{
    const int ta = (L.G.na + TX - 1) / TX, tb = (L.G.nb + TY - 1) / TY;
    const long long tiles = (long long)ta * tb;
    const int vgl = fwd_vg_log2(L.flags, L.nv), vg = 1 << vgl;
    long long max_groups = (0x7fffffffll / (TX * TY)) / (tiles * vg);
    if (max_groups < 1) max_groups = 1;
    const int chunk = (int)std::min<long long>(L.nv, max_groups * vg);
    for (int v0 = 0; v0 < L.nv; v0 += chunk) {
        const int nvc = std::min(chunk, L.nv - v0);
        const long long groups = (nvc + vg - 1) / vg;
        const long long blocks = (tiles * groups * vg + 7) / 8 * 8;
        hipLaunchKernelGGL((k_siddon<MODE, REF, TX, TY>), dim3((unsigned)blocks), dim3(TX, TY), 0, s, L, pv_in,
                           pv_out, counts, ta, tb, v0, nvc, vgl);
    }
}

template <int MODE, bool REF>
void launch_siddon(const Launch& L, const float* pv_in, float* pv_out, int32_t* counts, hipStream_t s)
{
    const unsigned t = (MODE == 0) ? (L.flags & NEFCT_F_FWD_TILE_MASK) : NEFCT_F_FWD_TILE_16x16;
    if (t == NEFCT_F_FWD_TILE_32x8)
        launch_siddon_tiles<MODE, REF, 32, 8>(L, pv_in, pv_out, counts, s);
    else if (t == NEFCT_F_FWD_TILE_8x32)
        launch_siddon_tiles<MODE, REF, 8, 32>(L, pv_in, pv_out, counts, s);
    else if (t == NEFCT_F_FWD_TILE_64x4)
        launch_siddon_tiles<MODE, REF, 64, 4>(L, pv_in, pv_out, counts, s);
    else
        launch_siddon_tiles<MODE, REF, FX, FY>(L, pv_in, pv_out, counts, s);
}

// ---------------------------------------------------------------------------
// Adjoint with LDS slab windows.
//
// A workgroup owns a tile of 64 columns x (4 * AR) rows of one view (each lane
// walks AR rays).  All its rays step through the same slab t together.  The
// segments of slab t are first summed into an LDS window covering the tile's
// footprint in that slab; after one barrier the window is flushed with one
// global float atomic per touched voxel (rows of the window are contiguous in
// the padded layout, so the flush is coalesced).  Two windows alternate, so
// one barrier per slab suffices.  A segment whose cell falls outside the
// window (loose bound) goes straight to a global atomic, so the window bound
// affects speed only, never the result.
// ---------------------------------------------------------------------------
constexpr int WIN_A = 128;
constexpr int WIN_B = 32;
constexpr int AR = 4;

__device__ __forceinline__ float wave_min_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, NEFCT_WAVE));
    return v;
}

__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, NEFCT_WAVE));
    return v;
}

template <bool REF, bool CHECKED>
__device__ __forceinline__ void adj_walk(const Ray (&r)[AR], const bool (&mine)[AR], const float (&sv)[AR], int lo,
                                         int hi, const Walk& w, __amdgpu_buffer_rsrc_t rs, float (*win)[WIN_B][WIN_A],
                                         const float (&bnd)[8], int wave, int lane)
{
    float a0[AR], b0[AR], fla0[AR], flb0[AR], fra0[AR], frb0[AR];
    int ca0[AR], cb0[AR];
#pragma unroll
    for (int j = 0; j < AR; ++j) {
        const float tf = (float)(lo - r[j].tc);
        a0[j] = fmaf(r[j].ka, tf, r[j].Ac);
        b0[j] = fmaf(r[j].kb, tf, r[j].Bc);
        fla0[j] = floorf(a0[j]);
        flb0[j] = floorf(b0[j]);
        ca0[j] = cell_floor(a0[j]);
        cb0[j] = cell_floor(b0[j]);
        fra0[j] = cell_frac(a0[j]);
        frb0[j] = cell_frac(b0[j]);
    }
    const float amn = bnd[0], amx = bnd[1], kamn = bnd[2], kamx = bnd[3];
    const float bmn = bnd[4], bmx = bnd[5], kbmn = bnd[6], kbmx = bnd[7];
    for (int t = lo; t < hi; ++t) {
        const float dt0 = (float)(t - lo), dt1 = dt0 + 1.0f;
        const float la_lo = amn + fminf(kamn * dt0, kamn * dt1), la_hi = amx + fmaxf(kamx * dt0, kamx * dt1);
        const float lb_lo = bmn + fminf(kbmn * dt0, kbmn * dt1), lb_hi = bmx + fmaxf(kbmx * dt0, kbmx * dt1);
        const int oa = __builtin_amdgcn_readfirstlane((int)floorf(la_lo) - 1);
        const int ob = __builtin_amdgcn_readfirstlane((int)floorf(lb_lo) - 1);
        const int span_a = __builtin_amdgcn_readfirstlane(min(max((int)floorf(la_hi) + 2 - oa, 0), WIN_A));
        const int span_b = __builtin_amdgcn_readfirstlane(min(max((int)floorf(lb_hi) + 2 - ob, 0), WIN_B));
        float(*cur)[WIN_A] = win[t & 1];
        const int tb = t * w.st + w.org;
#pragma unroll
        for (int j = 0; j < AR; ++j) {
            const float tf = (float)(t + 1 - r[j].tc);
            const float a1 = fmaf(r[j].ka, tf, r[j].Ac), b1 = fmaf(r[j].kb, tf, r[j].Bc);
            const float fla1 = floorf(a1), flb1 = floorf(b1);
            Seg s;
            int ca1 = 0, cb1 = 0;
            float fra1 = 0.0f, frb1 = 0.0f;
            if (REF) {
                slab_segments(a0[j], a1, fla0[j], fla1, b0[j], b1, flb0[j], flb1, w, s);
            } else {
                const float pia = r[j].pa * r[j].inv_ka, pib = r[j].pb * r[j].inv_kb;
                if (near_unit_slope(r[j])) {
                    slab_step(ca0[j], fra0[j], cb0[j], frb0[j], a1, b1, r[j], pia, pib, ca1, fra1, cb1, frb1, s);
                } else {
                    ca1 = cell_floor(a1);
                    cb1 = cell_floor(b1);
                    fra1 = cell_frac(a1);
                    frb1 = cell_frac(b1);
                    slab_exact<false>(ca0[j], ca1, fra0[j], cb0[j], cb1, frb0[j], r[j], pia, pib, s);
                }
            }
            const bool act = mine[j] && (t >= r[j].s0) && (t < r[j].s1) && s.ok;
            const int ca[3] = {s.a1, s.a2, s.a3};
            const int cb[3] = {s.b1, s.b2, s.b3};
            const float wk[3] = {s.w1, s.w2, s.w3};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                bool u = act && wk[k] > 0.0f;
                if (CHECKED || REF) u = u && cell_in(w, ca[k], cb[k]);
                if (u) {
                    const float v = wk[k] * sv[j];
                    const int la = ca[k] - oa, lb = cb[k] - ob;
                    if ((unsigned)la < (unsigned)span_a && (unsigned)lb < (unsigned)span_b)
                        atomicAdd(&cur[lb][la], v);
                    else
                        buf_atomic_add(rs, (ca[k] * w.sa + cb[k] * w.sb + tb) * 4, v);
                }
            }
            a0[j] = a1;
            b0[j] = b1;
            fla0[j] = fla1;
            flb0[j] = flb1;
            ca0[j] = ca1;
            cb0[j] = cb1;
            fra0[j] = fra1;
            frb0[j] = frb1;
        }
        __syncthreads();
        for (int lb = wave; lb < span_b; lb += BY) {
            for (int la = lane; la < span_a; la += BX) {
                const float v = cur[lb][la];
                if ((__float_as_uint(v) << 1) != 0u) {   // not +-0 (NaN included)
                    cur[lb][la] = 0.0f;
                    const int ca = oa + la, cb = ob + lb;
                    if (!(CHECKED || REF) || cell_in(w, ca, cb))
                        buf_atomic_add(rs, (ca * w.sa + cb * w.sb + tb) * 4, v);
                }
            }
        }
    }
}

template <bool REF>
__global__ __launch_bounds__(BX * BY) void k_siddon_adj(Launch P, const float* __restrict__ pv_in)
{
    __shared__ float win[2][WIN_B][WIN_A];
    __shared__ float redf[BY][8];
    __shared__ int redi[BY][2];
    const int lane = threadIdx.x, wave = threadIdx.y;
    const int tid = lane + BX * wave;
    {
        float* wf = &win[0][0][0];
        for (int i = tid; i < 2 * WIN_B * WIN_A; i += BX * BY) wf[i] = 0.0f;
    }
    const int ia = blockIdx.x * BX + lane;
    const int iv = blockIdx.z;
    const double c = P.trig[2 * iv], s = P.trig[2 * iv + 1];
    const double dz = view_lift(P, iv);
    Ray r[AR];
    bool in[AR];
    float sv[AR];
#pragma unroll
    for (int j = 0; j < AR; ++j) {
        const int ib = blockIdx.y * (BY * AR) + wave + BY * j;
        in[j] = ia < P.G.na && ib < P.G.nb;
        r[j].dom = -1;
        r[j].s0 = r[j].s1 = r[j].tc = 0;
        r[j].ka = r[j].kb = r[j].Ac = r[j].Bc = r[j].inv_ka = r[j].inv_kb = r[j].pa = r[j].pb = 0.0f;
        r[j].scale = 0.0f;
        sv[j] = 0.0f;
        if (in[j]) {
            setup_ray(P.V, P.G, c, s, ia, ib, P.flags, true, r[j], dz);
            const size_t id = (size_t)ia + (size_t)P.G.na * ((size_t)ib + (size_t)P.G.nb * (size_t)iv);
            sv[j] = pv_in[id] * r[j].scale;
        }
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (REF && d == 2) break;
        bool mine[AR];
        int lo = 0x7fffffff, hi = -0x7fffffff;
        float bnd[8] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
#pragma unroll
        for (int j = 0; j < AR; ++j) {
            mine[j] = in[j] && r[j].dom == d;
            if (mine[j]) {
                lo = min(lo, r[j].s0);
                hi = max(hi, r[j].s1);
            }
        }
        lo = wave_min_i32(lo);
        hi = wave_max_i32(hi);
        if (lane == 0) {
            redi[wave][0] = lo;
            redi[wave][1] = hi;
        }
        __syncthreads();
        lo = min(min(redi[0][0], redi[1][0]), min(redi[2][0], redi[3][0]));
        hi = max(max(redi[0][1], redi[1][1]), max(redi[2][1], redi[3][1]));
        lo = __builtin_amdgcn_readfirstlane(lo);
        hi = __builtin_amdgcn_readfirstlane(hi);
        __syncthreads();
        if (lo >= hi) continue;
#pragma unroll
        for (int j = 0; j < AR; ++j) {
            if (mine[j]) {
                const float tf = (float)(lo - r[j].tc);
                const float a = fmaf(r[j].ka, tf, r[j].Ac), b = fmaf(r[j].kb, tf, r[j].Bc);
                bnd[0] = fminf(bnd[0], a);
                bnd[1] = fmaxf(bnd[1], a);
                bnd[2] = fminf(bnd[2], r[j].ka);
                bnd[3] = fmaxf(bnd[3], r[j].ka);
                bnd[4] = fminf(bnd[4], b);
                bnd[5] = fmaxf(bnd[5], b);
                bnd[6] = fminf(bnd[6], r[j].kb);
                bnd[7] = fmaxf(bnd[7], r[j].kb);
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            bnd[q] = wave_min_f(bnd[q]);
            bnd[q + 1] = wave_max_f(bnd[q + 1]);
        }
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q) redf[wave][q] = bnd[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            float mn = redf[0][q], mx = redf[0][q + 1];
            for (int wv = 1; wv < BY; ++wv) {
                mn = fminf(mn, redf[wv][q]);
                mx = fmaxf(mx, redf[wv][q + 1]);
            }
            bnd[q] = mn;
            bnd[q + 1] = mx;
        }
        __syncthreads();
        // a bound that is not finite (no active ray had finite data) disables the window
        if (!(fabsf(bnd[0]) < 1e6f && fabsf(bnd[1]) < 1e6f && fabsf(bnd[4]) < 1e6f && fabsf(bnd[5]) < 1e6f)) {
            bnd[0] = bnd[1] = bnd[4] = bnd[5] = 0.0f;
            bnd[2] = bnd[3] = bnd[6] = bnd[7] = 0.0f;
        }
        const Walk w = walk_of(d, P.V, P.Ly);
        float* base = (d == 0) ? P.Ly.yf : P.Ly.xf;
        const uint32_t nbytes = (d == 0) ? P.Ly.yf_bytes : P.Ly.xf_bytes;
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(base, nbytes);
        if (d == 2)
            adj_walk<REF, true>(r, mine, sv, lo, hi, w, rs, win, bnd, wave, lane);
        else
            adj_walk<REF, false>(r, mine, sv, lo, hi, w, rs, win, bnd, wave, lane);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Adjoint owned by volume tiles (no global atomics).
//
// A workgroup owns a TX x TY x TZ block of voxels held in LDS.  For every view
// it enumerates the detector cells whose rays can reach the block (the
// block's shadow, from its 8 corners), sets each ray up exactly as the forward
// kernel does, walks only the slabs where the ray is within one cell of the
// block, and adds each segment that lands inside the block to LDS.  After
// the last view the block is written once.  Every (ray, slab, cell) weight is
// the forward kernel's own arithmetic, and every voxel belongs to one block,
// so the result is the transpose of the forward up to summation order.
//
// Accumulation (FIX = true, the default): each contribution w * pv * scale
// is converted to a signed fixed-point integer with a power-of-two scale S and
// added to a 64-bit integer in LDS (ds_add_u64).  S is the largest power of
// two with |contribution| * S <= 2^30 for every contribution of the call
// (from max |pv| over the sinogram and a geometric bound on the per-slab ray
// length), so the rounding step is ~1e-9 of the largest contribution, and
// integer sums do not depend on the order of the additions: the result is
// bitwise reproducible.  LDS float atomics (ds_add_f32, FIX = false) measure
// several times slower than integer ones on gfx950, which is why this exists.
// A non-finite projection value makes the affected blocks NaN.
// ---------------------------------------------------------------------------
template <int TX, int TY, int TZ, typename T, bool HALO_ = false>
struct TileAcc {
    // odd row and plane strides (in elements): lanes on consecutive y rows or
    // z planes hit distinct banks
    static constexpr bool HALO = HALO_;
    static constexpr int SX = TX + 1;
    static constexpr int SXY0 = (HALO ? TY + 1 : TY) * SX;
    static constexpr int SXY = SXY0 + (SXY0 % 2 == 0 ? 1 : 0);
    // With HALO, every cell (x, y, z) with x in [-1, TX], y in [-1, TY], z in
    // [-1, TZ] has a word of its own or shares one with another such halo cell
    // (x = TX of row y is x = -1 of row y + 1; the planes are at least TY + 1
    // rows apart), never with a block cell.  BASE is cell (0, 0, 0).
    static constexpr int BASE = HALO ? SXY + SX + 1 : 0;
    static constexpr int NEL = HALO ? BASE + TZ * SXY + TY * SX + TX + 1 : TZ * SXY;
    T v[NEL];
    __device__ T& at(int x, int y, int z) { return v[BASE + z * SXY + y * SX + x]; }
};

template <bool FIX>
using acc_t = typename std::conditional<FIX, unsigned long long, float>::type;

// Exponent e of the power-of-two fixed-point scale 2^e: the largest e with
// max |projection| * sbound * 2^e <= 2^30, from the sinogram's max |value|
// (as float bits) and the bound on the per-slab ray length.  Computed in FP64:
// for tiny sinograms 2^e lies beyond the FP32 range (2^30 / 1e-32 ~ 2^136),
// so contributions are scaled by ldexpf(x, e), never by a float 2^e.
// FIX_POISON marks non-finite input; an all-zero sinogram gives 0.
constexpr int FIX_POISON = -0x40000000;
__device__ __forceinline__ int fixed_exp(unsigned maxbits, float sbound)
{
    // Inf or NaN somewhere (tested on the bits: NaN and Inf sort above 3e38)
    if (maxbits > __builtin_bit_cast(unsigned, 3.0e38f)) return FIX_POISON;
    const double big = (double)__uint_as_float(maxbits) * (double)sbound;
    if (!(big > 0.0)) return 0;                 // all zero
    int e;
    frexp(1073741824.0 / big, &e);              // 2^30 / big = m 2^e, m in [0.5, 1)
    return e - 1;                               // 2^(e-1) <= 2^30 / big
}

// Adds contribution v to the LDS word at byte_off.  Fixed point: v already
// carries the scale S (the ray's value is pre-multiplied by S once), and the
// conversion truncates toward zero: the per-contribution error is below one
// unit of 2^-30 of the largest contribution, independent of order.
template <bool FIX>
__device__ __forceinline__ void lds_deposit(char* base, int byte_off, float v)
{
    if (FIX) {
        const long long q = (long long)(int)v;
        atomicAdd(reinterpret_cast<unsigned long long*>(base + byte_off), (unsigned long long)q);
    } else {
        atomicAdd(reinterpret_cast<float*>(base + byte_off), v);
    }
}

// Fixed-point deposit of quantum q with sign word hi (0 or -1); with hi = -1
// q must be nonzero (a zero quantum would carry the sign word alone).
__device__ __forceinline__ void lds_deposit_q(char* base, int byte_off, int q, int hi)
{
    const unsigned long long v = ((unsigned long long)(unsigned)hi << 32) | (unsigned)q;
    atomicAdd(reinterpret_cast<unsigned long long*>(base + byte_off), v);
}

// Fixed-point deposit of any quantum q (sign-extended to 64 bits).
__device__ __forceinline__ void lds_deposit_s(char* base, int byte_off, int q)
{
    atomicAdd(reinterpret_cast<unsigned long long*>(base + byte_off), (unsigned long long)(long long)q);
}

// Detector cells whose rays can meet the voxel box [lo, hi) at view (c, s),
// as (ia0, ia1, ib0, ib1) (empty when ia1 < ia0).  Computed by 8 lanes
// together: lane q (0..7) of the group takes corner q.
__device__ __forceinline__ int4 shadow_of_box(const VolDev& V, const DetDev& G, double c, double s, const int lo[3],
                                              const int hi[3], bool tight, int q, double dz = 0.0)
{
    // Fractional detector indices of the 8 corners (flat: a = t SDD / U,
    // b = z SDD / U; cylindrical: a = atan2(t, U), b = z SDD / hypot(U, t)),
    // with the margin chosen below.  The 8 lanes reduce by butterfly, so
    // every lane of the group ends with the same bounds.
    const double x = V.lo[0] + V.vs[0] * ((q & 1) ? hi[0] : lo[0]);
    const double y = V.lo[1] + V.vs[1] * ((q & 2) ? hi[1] : lo[1]);
    const double z = V.lo[2] + V.vs[2] * ((q & 4) ? hi[2] : lo[2]) - dz;   // relative to the lifted source
    const double U = G.sid + x * c + y * s;
    const double t = -x * s + y * c;
    int behind = (U <= 1e-6 * G.sid) ? 1 : 0;
    double a, b;
    if (G.cyl) {
        a = atan2(t, U);
        b = z * G.sad / sqrt(U * U + t * t);
    } else {
        const double iu = G.sad / U;
        a = t * iu;
        b = z * iu;
    }
    float fa_mn = (float)((a + G.na * G.da * 0.5) / G.da - G.ai);
    float fb_mn = (float)((b + G.nb * G.db * 0.5) / G.db - G.bi);
    float fa_mx = fa_mn, fb_mx = fb_mn;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        fa_mn = fminf(fa_mn, __shfl_xor(fa_mn, m));
        fa_mx = fmaxf(fa_mx, __shfl_xor(fa_mx, m));
        fb_mn = fminf(fb_mn, __shfl_xor(fb_mn, m));
        fb_mx = fmaxf(fb_mx, __shfl_xor(fb_mx, m));
        behind |= __shfl_xor(behind, m);
    }
    if (behind)   // block reaches behind the source: take the whole detector
        return make_int4(0, G.na - 1, 0, G.nb - 1);
    if (tight && !G.cyl) {
        // Flat detector, exact rule: the block's projection is the convex
        // hull of its corners' (a linear-fractional map), and cell i's ray
        // goes through fractional index i, so the cells whose rays meet the
        // block are ceil(min) .. floor(max).  The margin (1/8 cell) covers
        // the FP32 rounding of these indices and of the walk's coordinates
        // many times over.
        constexpr float M = 0.125f;
        return make_int4(max(0, ceil_clamp(fa_mn - M, -2, G.na + 1)), min(G.na - 1, floor_clamp(fa_mx + M, -2, G.na + 1)),
                         max(0, ceil_clamp(fb_mn - M, -2, G.nb + 1)), min(G.nb - 1, floor_clamp(fb_mx + M, -2, G.nb + 1)));
    }
    // one cell of margin: on a cylindrical detector the row coordinate of an
    // edge point can exceed its corners'; the reference rule can credit a
    // cell next to its exact segment
    return make_int4(max(0, floor_clamp(fa_mn, -2, G.na + 1) - 1), min(G.na - 1, ceil_clamp(fa_mx, -2, G.na + 1) + 1),
                     max(0, floor_clamp(fb_mn, -2, G.nb + 1) - 1), min(G.nb - 1, ceil_clamp(fb_mx, -2, G.nb + 1) + 1));
}

// A block's shadows (and the views' sources in voxel units) are computed SHV
// views at a time into LDS tables by all of the workgroup's lanes (8 per
// view), instead of by every wavefront for every view.
constexpr int SHV = 128;

// A block's accumulator carries the one-cell halo (walks without range tests)
// when it fits in the 160 KB of LDS beside the per-view tables (shadow,
// source, divisor: 44 B per view) and the axis records.
constexpr size_t LDS_BYTES = 160 * 1024;
constexpr size_t BLOCK_TABLE_BYTES = SHV * (sizeof(int4) + 3 * sizeof(double) + sizeof(unsigned)) + 1024;
template <int TX, int TY, int TZ, bool FIX>
constexpr bool halo_fits()
{
    return sizeof(TileAcc<TX, TY, TZ, acc_t<FIX>, true>) + BLOCK_TABLE_BYTES <= LDS_BYTES;
}
template <int TX, int TY, int TZ, bool FIX>
using BlockAcc = TileAcc<TX, TY, TZ, acc_t<FIX>, halo_fits<TX, TY, TZ, FIX>()>;


// What a ray walking axis m needs of its block: the slab range on m, the
// block's first centred cell and extent on the transverse axes a and b, the
// MARGIN-widened box in centred coordinates, and the LDS byte strides (one
// record per m, read by each lane for its ray's axis; selecting them from the
// block bounds per lane costs ~40 instructions per ray).
struct BlockAxis {
    int b0m, b1m;
    int la, lb;
    unsigned ea, eb;
    float alo, ahi, blo, bhi;
    int Sa, Sb, St;
    int row0;        // LDS byte offset of slab 0's centred cell (0, 0): row = t St + row0
    // the block's transverse range widened by the halo cell on each side:
    // positions a with floor(a) in [la - 1, la + ea] are those in [hla, hha)
    float hla, hha, hlb, hhb;
    int pad_[2];
};

template <bool REF, bool FIX, int TX, int TY, int TZ, int NW, int CS>
__device__ __forceinline__ void adj_tile_block(const Launch& P, const float* __restrict__ pv, float* __restrict__ image,
                                               int bx, int by, int bz, BlockAcc<TX, TY, TZ, FIX>& acc,
                                               int4* __restrict__ shadows, int es, bool accumulate)
{
    using T = acc_t<FIX>;
    using Acc = BlockAcc<TX, TY, TZ, FIX>;
    constexpr int SX = Acc::SX;
    constexpr int SXY = Acc::SXY;
    constexpr bool HALO = Acc::HALO;
    constexpr int ES = (int)sizeof(T);
    const int tid = threadIdx.x;
    T* af = &acc.v[Acc::BASE];   // cell (0, 0, 0); halo cells lie at negative offsets and beyond TZ planes
    __syncthreads();   // the previous block's write-out has read acc
    for (int i = tid; i < TZ * SXY; i += 64 * NW) af[i] = (T)0;
    const int nx = P.V.n[0], ny = P.V.n[1], nz = P.V.n[2];
    const int b0[3] = {bx * TX, by * TY, bz * TZ};
    const int ext[3] = {min(TX, nx - b0[0]), min(TY, ny - b0[1]), min(TZ, nz - b0[2])};
    const int b1[3] = {b0[0] + ext[0], b0[1] + ext[1], b0[2] + ext[2]};
    constexpr float MARGIN = 1.0f / 64.0f;   // see the range tightening below
    __shared__ BlockAxis bax[3];
    __shared__ AxisGeom vax[3];          // volume part of each walk axis's AxisGeom
    __shared__ double src_tab[SHV][3];   // source in voxel units per view of the chunk
    __shared__ unsigned mh_tab[SHV];     // ceil(2^32 / H) of each view's shadow height H (cell dealing)
    if (tid < 3) {
        double zero[3] = {0.0, 0.0, 0.0};
        vax[tid] = axis_geom(P.V, zero, tid);
    }
    if (tid < 3) {
        const int m = tid, a = (m == 0) ? 1 : 0, b = (m == 2) ? 1 : 2;
        BlockAxis X;
        X.b0m = b0[m];
        X.b1m = b1[m];
        X.la = b0[a] - (P.V.n[a] >> 1);
        X.lb = b0[b] - (P.V.n[b] >> 1);
        X.ea = (unsigned)ext[a];
        X.eb = (unsigned)ext[b];
        X.alo = (float)X.la - MARGIN;
        X.ahi = (float)(X.la + ext[a]) + MARGIN;
        X.blo = (float)X.lb - MARGIN;
        X.bhi = (float)(X.lb + ext[b]) + MARGIN;
        X.Sa = ES * ((m == 0) ? SX : 1);
        X.Sb = ES * ((m == 2) ? SX : SXY);
        X.St = ES * ((m == 0) ? 1 : ((m == 1) ? SX : SXY));
        X.row0 = -X.b0m * X.St - X.la * X.Sa - X.lb * X.Sb;
        X.hla = (float)(X.la - 1);
        X.hha = (float)(X.la + ext[a] + 1);
        X.hlb = (float)(X.lb - 1);
        X.hhb = (float)(X.lb + ext[b] + 1);
        X.pad_[0] = X.pad_[1] = 0;
        bax[m] = X;
    }
    __syncthreads();
    for (int iv = 0; iv < P.nv; ++iv) {
        const int slot = iv % SHV;
        if (slot == 0) {
            __syncthreads();   // every wavefront is done with the previous views' entries
            const int n = min(SHV, P.nv - iv);
            for (int e = tid >> 3; e < n; e += 8 * NW) {   // the 8 lanes of a group share e
                const int jv = iv + e;
                const double cj = P.trig[2 * jv], sj = P.trig[2 * jv + 1];
                const double dzj = view_lift(P, jv);
                const int4 sh = shadow_of_box(P.V, P.G, cj, sj, b0, b1, !REF, tid & 7, dzj);
                if ((tid & 7) == 0) {
                    shadows[e] = sh;
                    // a 64-bit division once per block and view (not in every
                    // wavefront for every view)
                    const unsigned Hh = (sh.w >= sh.z) ? (unsigned)(sh.w - sh.z + 1) : 1u;
                    mh_tab[e] = (Hh > 1u) ? (unsigned)((0x100000000ull + Hh - 1) / Hh) : 0u;
                    source_voxels(P.V, P.G, cj, sj, src_tab[e], dzj);
                }
            }
            __syncthreads();
        }
        const int4 sh = shadows[slot];
        const int ia0 = __builtin_amdgcn_readfirstlane(sh.x), ia1 = __builtin_amdgcn_readfirstlane(sh.y);
        const int ib0 = __builtin_amdgcn_readfirstlane(sh.z), ib1 = __builtin_amdgcn_readfirstlane(sh.w);
        if (ia1 < ia0 || ib1 < ib0) continue;
        const double c = P.trig[2 * iv], s = P.trig[2 * iv + 1];
        // The shadow's W x H detector cells are numbered column by column
        // and dealt out densely: wavefront w takes cells w*64 + lane +
        // j*64*NW, so no lane idles for rounding the shadow up to whole
        // 64-cell chunks.  A wavefront's lanes are then neighbouring rows of
        // one or two detector columns: rays of one column share their path in
        // the plane of the source trajectory (flat and cylindrical detectors
        // alike), so they cross the block over nearly the same slab range and
        // finish the walk together (numbering row by row mixes rays through
        // the block's middle with rays grazing its edges in one wavefront).
        // CS > 1 interleaves rows instead: lane l of a 64*CS-cell run takes
        // row CS*l + phase.
        const int wv = tid >> 6, ln = tid & 63;
        const unsigned W = (unsigned)(ia1 - ia0 + 1), H = (unsigned)(ib1 - ib0 + 1);
        const unsigned ncell = W * H;
        const unsigned mH = __builtin_amdgcn_readfirstlane(mh_tab[slot]);   // ceil(2^32 / H), 0 for H = 1
        const unsigned groups = (ncell + 64u * CS - 1) / (64u * CS) * CS;   // 64-cell groups to deal out
        const int nk = (int)((groups + NW - 1) / NW);
        auto cell_of = [&](int k, int& ia, int& ib) {
            unsigned e = (unsigned)(wv * 64) + (unsigned)k * (64u * NW);
            if (CS > 1) {
                const unsigned run = e / (64u * CS) * (64u * CS);   // e is a multiple of 64
                e = run + (unsigned)(CS * ln) + (e - run) / 64u;
            } else {
                e += (unsigned)ln;
            }
            // e / H by multiplication (the quotient may come out one too large)
            unsigned q = (H > 1u) ? __umulhi(e, mH) : e;
            unsigned rr = e - q * H;
            if ((int)rr < 0) { q -= 1; rr += H; }
            ia = ia0 + (int)q;
            ib = ib0 + (int)rr;
            return e < ncell;
        };
        // the view's sinogram; a cell's offset in it fits 32 bits (na nb < 2^31)
        const float* __restrict__ pvv = pv + (size_t)P.G.na * (size_t)P.G.nb * (size_t)iv;
        int ia_n = 0, ib_n = 0;
        bool ok_n = nk > 0 && cell_of(0, ia_n, ib_n);
        float pv_n = ok_n ? pvv[(unsigned)(ia_n + P.G.na * ib_n)] : 0.0f;
        for (int k = 0; k < nk; ++k) {
            const int ia = ia_n, ib = ib_n;
            const bool ok = ok_n;
            const float pval = pv_n;
            ok_n = k + 1 < nk && cell_of(k + 1, ia_n, ib_n);
            if (ok_n) pv_n = pvv[(unsigned)(ia_n + P.G.na * ib_n)];
            if (!ok) continue;
            Ray r;
            // the set-up's per-axis values from the LDS tables (the same
            // values setup_ray selects from the volume record)
            setup_ray_g(P.V, P.G, c, s, ia, ib, P.flags, true,
                        [&](int m) {
                            AxisGeom g = vax[m];
                            g.uSm = src_tab[slot][m];
                            g.uSa = src_tab[slot][(m == 0) ? 1 : 0];
                            g.uSb = src_tab[slot][(m == 2) ? 1 : 2];
                            return g;
                        },
                        r);
            if (r.dom < 0) continue;
            const int m = r.dom;
            const BlockAxis& X = bax[m];
            // slabs of the block on the walk axis, within the ray's own range
            int t0 = max(r.s0, X.b0m), t1 = min(r.s1, X.b1m);
            if (t0 >= t1) continue;
            // FP32 coordinates of the ray at the block's slab faces decide the
            // transverse overlap, as the walk will compute them.  The box is
            // the block widened by MARGIN cells: a slab deposits only into
            // cells between floor(a(t)) and floor(a(t+1)), so it matters only
            // if the walk's position there reaches the block's range; the
            // walk's FP32 positions are within 2^-24 * 512 cells of the exact
            // line and the slab estimates below within 600 * 2^-23 * |1/k|
            // slabs, both far inside MARGIN (and its MARGIN / |k| slabs).
            // Slab t (positions [t, t + 1]) meets the widened range [u, v]
            // iff floor(u) <= t < ceil(v) up to ties, where the slab only
            // touches a widened face.  The reference rule can credit a
            // neighbour of its exact cell, so it keeps one slab of slack on
            // each side.
            constexpr int SLACK = REF ? 1 : 0;
            const float fa0 = fmaf(r.ka, (float)(t0 - r.tc), r.Ac), fa1 = fmaf(r.ka, (float)(t1 - r.tc), r.Ac);
            const float fb0 = fmaf(r.kb, (float)(t0 - r.tc), r.Bc), fb1 = fmaf(r.kb, (float)(t1 - r.tc), r.Bc);
            const float alo = X.alo, ahi = X.ahi, blo = X.blo, bhi = X.bhi;
            const int bm0 = X.b0m, bm1 = X.b1m;
            if (fmaxf(fa0, fa1) < alo || fminf(fa0, fa1) > ahi || fmaxf(fb0, fb1) < blo || fminf(fb0, fb1) > bhi)
                continue;
            // tighten [t0, t1) to the slabs where the ray is within the margin
            // box, with SLACK slabs on each side (the clamps keep the
            // result while bounding the conversion: t0 and t1 already lie in
            // [b0[m], b1[m]]).  A slope below 1e-30 leaves the range as is;
            // otherwise 1/k < ~1e30 and the estimates are finite.
            if (fabsf(r.ka) >= 1e-30f) {
                float ta = (alo - r.Ac) * r.inv_ka + (float)r.tc, tb = (ahi - r.Ac) * r.inv_ka + (float)r.tc;
                if (ta > tb) { const float q = ta; ta = tb; tb = q; }
                t0 = max(t0, floor_clamp(ta, bm0, bm1 + 2) - SLACK);
                t1 = min(t1, ceil_clamp(tb, bm0 - 2, bm1) + SLACK);
            }
            if (fabsf(r.kb) >= 1e-30f) {
                float ta = (blo - r.Bc) * r.inv_kb + (float)r.tc, tb = (bhi - r.Bc) * r.inv_kb + (float)r.tc;
                if (ta > tb) { const float q = ta; ta = tb; tb = q; }
                t0 = max(t0, floor_clamp(ta, bm0, bm1 + 2) - SLACK);
                t1 = min(t1, ceil_clamp(tb, bm0 - 2, bm1) + SLACK);
            }
            if (t0 >= t1) continue;
            const float sv = FIX ? ldexpf(pval * r.scale, es) : pval * r.scale;   // fixed point: scaled once per ray
            const Walk w = walk_of(m, P.V, P.Ly);
            // LDS byte offset of centred transverse cell (ca, cb) at slab t:
            // (ca Sa + cb Sb + t St + base) * ES, with the strides of the walk
            // axis; [t0, t1) lies inside the block, so only ca and cb are
            // range-checked
            const int Sa = X.Sa, Sb = X.Sb, St = X.St;
            const int la = X.la, lb = X.lb;     // centred cell of the block's first a / b index
            const unsigned ea = X.ea, eb = X.eb;
            char* const ab = reinterpret_cast<char*>(af);
            int row = t0 * St + X.row0;
            float tf = (float)(t0 - r.tc);
            float a0 = fmaf(r.ka, tf, r.Ac), bb0 = fmaf(r.kb, tf, r.Bc);
            float fla0 = floorf(a0), flb0 = floorf(bb0);
            if (!REF) {
                // Exact rule (slab_exact, the forward's arithmetic) with the
                // slab's entry cell carried from the previous slab: its index,
                // fraction, range test and LDS offset E are computed once per
                // slab.  The exit cell's offset is E + dA + dB, the middle
                // piece's E + dA or E + dB, and the next entry is exit + St.
                // near_unit_slope rays step their cells with the crossings
                // (slab_step), as the forward walks them.
                const float pia = r.pa * r.inv_ka, pib = r.pb * r.inv_kb;
                const int qs = FIX ? (int)sv : 0;               // |sv| <= 2^30
                const int hi = (FIX && sv < 0.0f) ? -1 : 0;
                // the loop ends on the slab coordinate itself (exact small
                // integers in FP32), two slabs per trip
                const float tend = (float)(t1 - r.tc);
                // The walk visits the positions a(t0), a(t0 + 1), ..., a(t1)
                // (one FP32 fma of an exact slab coordinate each, monotone in
                // t) and deposits into their floors (the crossing-stepped walk
                // of near_unit_slope rays aside).  A ray whose end positions
                // lie in the halo-widened range deposits only into block and
                // halo cells; a wavefront of such rays walks without range
                // tests, its out-of-block deposits landing in halo words that
                // are never read.
                bool in_halo = false;
                if constexpr (HALO) {
                    const float a_e = fmaf(r.ka, tend, r.Ac), b_e = fmaf(r.kb, tend, r.Bc);
                    const float hla = X.hla, hha = X.hha, hlb = X.hlb, hhb = X.hhb;
                    // bitwise and: all four tests at once, no branch per test
                    in_halo = (fminf(a0, a_e) >= hla) & (fmaxf(a0, a_e) < hha) & (fminf(bb0, b_e) >= hlb) &
                              (fmaxf(bb0, b_e) < hhb);
                }
                auto walk = [&](auto step_tag, auto nonneg_tag, auto check_tag) {
                    constexpr bool STEP = decltype(step_tag)::value;
                    constexpr bool NONNEG = decltype(nonneg_tag)::value;
                    constexpr bool CHECK = decltype(check_tag)::value;
                    float tw = tf;
                    int ca0 = cell_floor(a0), cb0 = cell_floor(bb0);
                    float fra0 = cell_frac(a0), frb0 = cell_frac(bb0);
                    int A0 = __mul24(ca0, Sa), B0 = __mul24(cb0, Sb);
                    int E = A0 + B0 + row;
                    bool ia0 = !CHECK || (unsigned)(ca0 - la) < ea, ib0 = !CHECK || (unsigned)(cb0 - lb) < eb;
                    auto slab = [&]() {
                        tw += 1.0f;
                        const float a1 = fmaf(r.ka, tw, r.Ac), bb1 = fmaf(r.kb, tw, r.Bc);
                        int ca1, cb1;
                        float fra1, frb1, fa, fb;
                        if constexpr (STEP) {
                            fa = (r.ka != 0.0f) ? cross_frac(fra0, r.inv_ka, pia) : 1.0f;
                            fb = (r.kb != 0.0f) ? cross_frac(frb0, r.inv_kb, pib) : 1.0f;
                            ca1 = ca0 + ((fa < 1.0f) ? ((r.ka > 0.0f) ? 1 : -1) : 0);
                            cb1 = cb0 + ((fb < 1.0f) ? ((r.kb > 0.0f) ? 1 : -1) : 0);
                            fra1 = a1 - (float)ca1;
                            frb1 = bb1 - (float)cb1;
                        } else {
                            ca1 = cell_floor(a1);
                            cb1 = cell_floor(bb1);
                            fra1 = cell_frac(a1);
                            frb1 = cell_frac(bb1);
                            fa = cross_frac(fra0, r.inv_ka, pia);
                            fb = cross_frac(frb0, r.inv_kb, pib);
                        }
                        const bool afirst = fa < fb;
                        const float f1 = afirst ? fa : fb, f2 = afirst ? fb : fa;
                        const int A1 = __mul24(ca1, Sa), B1 = __mul24(cb1, Sb);
                        const int dA = A1 - A0, dB = B1 - B0;
                        const int X = E + dA + dB;
                        const bool ia1 = !CHECK || (unsigned)(ca1 - la) < ea, ib1 = !CHECK || (unsigned)(cb1 - lb) < eb;
                        const bool i2 = (afirst & ia1 & ib0) | (!afirst & ia0 & ib1);
                        if constexpr (FIX) {
                            // the three pieces as differences of the truncated
                            // prefix sums f1 sv, f2 sv, sv: they add up to the
                            // slab's whole contribution exactly, and all share
                            // the ray's sign.  The last piece is nonzero only
                            // when the slab crosses both transverse boundaries:
                            // it is skipped when zero and its high word is the
                            // ray's constant.  The first and middle pieces are
                            // added even when zero (cheaper than the test): in
                            // a wavefront of non-negative rays with the high
                            // word 0, otherwise sign-extended (a zero quantum
                            // of a negative ray must not carry the sign word).
                            const int q1 = (int)(f1 * sv), q2 = (int)(f2 * sv);
                            const int d2 = q2 - q1, d3 = qs - q2;
                            if constexpr (NONNEG) {
                                if (ia0 && ib0) lds_deposit_q(ab, E, q1, 0);
                                if (i2) lds_deposit_q(ab, E + (afirst ? dA : dB), d2, 0);
                                if (ia1 && ib1 && d3 != 0) lds_deposit_q(ab, X, d3, 0);
                            } else {
                                if (ia0 && ib0) lds_deposit_s(ab, E, q1);
                                if (i2) lds_deposit_s(ab, E + (afirst ? dA : dB), d2);
                                if (ia1 && ib1 && d3 != 0) lds_deposit_q(ab, X, d3, hi);
                            }
                        } else {
                            const float w1 = f1, w2 = f2 - f1, w3 = 1.0f - f2;
                            if (ia0 && ib0 && w1 > 0.0f) lds_deposit<FIX>(ab, E, w1 * sv);
                            if (i2 && w2 > 0.0f) lds_deposit<FIX>(ab, E + (afirst ? dA : dB), w2 * sv);
                            if (ia1 && ib1 && w3 > 0.0f) lds_deposit<FIX>(ab, X, w3 * sv);
                        }
                        ca0 = ca1;
                        cb0 = cb1;
                        fra0 = fra1;
                        frb0 = frb1;
                        A0 = A1;
                        B0 = B1;
                        E = X + St;
                        ia0 = ia1;
                        ib0 = ib1;
                    };
                    for (;;) {
                        slab();
                        if (tw >= tend) break;
                        slab();
                        if (tw >= tend) break;
                    }
                };
                auto walk_slope = [&](auto nonneg_tag) {
                    if (near_unit_slope(r)) {
                        walk(std::true_type{}, nonneg_tag, std::true_type{});
                    } else {
                        if constexpr (HALO) {
                            if (__all(in_halo)) {   // wave-uniform
                                walk(std::false_type{}, nonneg_tag, std::false_type{});
                                return;
                            }
                        }
                        walk(std::false_type{}, nonneg_tag, std::true_type{});
                    }
                };
                // wave-uniform: every walking ray of the wavefront non-negative
                if (FIX && __all(hi == 0))
                    walk_slope(std::true_type{});
                else
                    walk_slope(std::false_type{});
                continue;
            }
            for (int t = t0; t < t1; ++t) {
                tf += 1.0f;
                const float a1 = fmaf(r.ka, tf, r.Ac), bb1 = fmaf(r.kb, tf, r.Bc);
                const float fla1 = floorf(a1), flb1 = floorf(bb1);
                Seg sg;
                slab_segments(a0, a1, fla0, fla1, bb0, bb1, flb0, flb1, w, sg);
                if (sg.ok) {
                    const int ca[3] = {sg.a1, sg.a2, sg.a3};
                    const int cb[3] = {sg.b1, sg.b2, sg.b3};
                    const float wk[3] = {sg.w1, sg.w2, sg.w3};
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        if (wk[q] > 0.0f && (unsigned)(ca[q] - la) < ea && (unsigned)(cb[q] - lb) < eb)
                            lds_deposit<FIX>(ab, ca[q] * Sa + cb[q] * Sb + row, wk[q] * sv);
                    }
                }
                a0 = a1;
                bb0 = bb1;
                fla0 = fla1;
                flb0 = flb1;
                row += St;
            }
        }
    }
    __syncthreads();
    // the scale is 2^es, or FIX_POISON for a non-finite sinogram: then every
    // voxel is written as a NaN bit pattern (an integer store)
    const bool poisoned = FIX && es == FIX_POISON;
    const double inv_s = (FIX && !poisoned) ? ldexp(1.0, -es) : 1.0;
    for (int i = tid; i < TX * TY * TZ; i += 64 * NW) {
        const int lx = i % TX, ly = (i / TX) % TY, lz = i / (TX * TY);
        if (lx < ext[0] && ly < ext[1] && lz < ext[2]) {
            const size_t o = (size_t)(b0[0] + lx) + (size_t)nx * ((size_t)(b0[1] + ly) + (size_t)ny * (size_t)(b0[2] + lz));
            if (poisoned) {
                reinterpret_cast<unsigned*>(image)[o] = 0x7fc00000u;
                continue;
            }
            float v;
            if (FIX)
                v = (float)((double)(long long)acc.at(lx, ly, lz) * inv_s);
            else
                v = (float)acc.at(lx, ly, lz);
            image[o] = accumulate ? image[o] + v : v;
        }
    }
}

template <bool REF, bool FIX, int TX, int TY, int TZ, int NW, int CS>
__global__ __launch_bounds__(64 * NW, 4) void k_siddon_adj_tile(Launch P, const float* __restrict__ pv,
                                                                float* __restrict__ image,
                                                                const unsigned* __restrict__ maxbits, float sbound)
{
    __shared__ BlockAcc<TX, TY, TZ, FIX> acc;
    __shared__ int4 shadows[SHV];
    const int es = FIX ? fixed_exp(*maxbits, sbound) : 0;
    adj_tile_block<REF, FIX, TX, TY, TZ, NW, CS>(P, pv, image, blockIdx.x, blockIdx.y, blockIdx.z, acc, shadows, es,
                                                 false);
}

// Launch of one tile-adjoint variant.
template <int TX, int TY, int TZ, int NW, int CS>
void launch_tile(bool ref, bool fix, const Launch& L, const float* pv, float* image, const unsigned* maxbits,
                 float sbound, hipStream_t s)
{
    const dim3 g((L.V.n[0] + TX - 1) / TX, (L.V.n[1] + TY - 1) / TY, (L.V.n[2] + TZ - 1) / TZ);
    const dim3 b(64 * NW);
    if (ref && fix)
        hipLaunchKernelGGL((k_siddon_adj_tile<true, true, TX, TY, TZ, NW, CS>), g, b, 0, s, L, pv, image, maxbits,
                           sbound);
    else if (ref)
        hipLaunchKernelGGL((k_siddon_adj_tile<true, false, TX, TY, TZ, NW, CS>), g, b, 0, s, L, pv, image, maxbits,
                           sbound);
    else if (fix)
        hipLaunchKernelGGL((k_siddon_adj_tile<false, true, TX, TY, TZ, NW, CS>), g, b, 0, s, L, pv, image, maxbits,
                           sbound);
    else
        hipLaunchKernelGGL((k_siddon_adj_tile<false, false, TX, TY, TZ, NW, CS>), g, b, 0, s, L, pv, image, maxbits,
                           sbound);
}

// max |x| over n values, as float bits, into *out (which the caller zeroes);
// NaN compares above Inf in this order
__global__ __launch_bounds__(256) void k_absmax(const float* __restrict__ x, size_t n, unsigned* __restrict__ out)
{
    unsigned m = 0u;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        m = max(m, __float_as_uint(x[i]) & 0x7fffffffu);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_view_trig(const float* __restrict__ angles, int nv, double* __restrict__ trig)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const double a = (double)angles[i];
    trig[2 * i] = cos(a);
    trig[2 * i + 1] = sin(a);
}

// Fill [p, p+n) with zeros (a kernel, not hipMemsetAsync, so it records
// cleanly into HIP graphs).
__global__ void k_zero(float* __restrict__ p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0f;
}

// x-fastest padded copy: out[(px) + NXP*(iy + ny*pz)]
__global__ void k_pad_x(const float* __restrict__ img, int nx, int ny, int nz, int NXP, float* __restrict__ out)
{
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    const int pz = blockIdx.z;
    if (px >= NXP) return;
    const int ix = px - PAD, iz = pz - PAD;
    float v = 0.0f;
    if (ix >= 0 && ix < nx && iz >= 0 && iz < nz) v = img[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)];
    out[(size_t)px + (size_t)NXP * ((size_t)iy + (size_t)ny * pz)] = v;
}

// y-fastest padded copy through a 64x64 LDS tile: out[(py) + NYP*(ix + nx*pz)]
__global__ __launch_bounds__(256) void k_pad_y(const float* __restrict__ img, int nx, int ny, int nz, int NYP,
                                               float* __restrict__ out)
{
    __shared__ float tile[64][65];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 64, pz = blockIdx.z;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int iz = pz - PAD;
    const bool zin = iz >= 0 && iz < nz;
#pragma unroll 4
    for (int r = ty; r < 64; r += 4) {
        const int iy = y0 + r - PAD, ix = x0 + tx;
        float v = 0.0f;
        if (zin && iy >= 0 && iy < ny && ix < nx) v = img[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)];
        tile[r][tx] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int r = ty; r < 64; r += 4) {
        const int ix = x0 + r, py = y0 + tx;
        if (ix < nx && py < NYP) out[(size_t)py + (size_t)NYP * ((size_t)ix + (size_t)nx * pz)] = tile[tx][r];
    }
}

// image[ix + nx*(iy + ny*iz)] = XF + YF accumulators (adjoint epilogue), tiled.
__global__ __launch_bounds__(256) void k_combine(const float* __restrict__ xf, const float* __restrict__ yf,
                                                 int nx, int ny, int nz, int NXP, int NYP,
                                                 float* __restrict__ img)
{
    __shared__ float tile[64][65];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 64, iz = blockIdx.z;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int pz = iz + PAD;
    // load y-fastest tile: rows ix, columns iy
#pragma unroll 4
    for (int r = ty; r < 64; r += 4) {
        const int ix = x0 + r, iy = y0 + tx;
        float v = 0.0f;
        if (ix < nx && iy < ny) v = yf[(size_t)(iy + PAD) + (size_t)NYP * ((size_t)ix + (size_t)nx * pz)];
        tile[r][tx] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int r = ty; r < 64; r += 4) {
        const int iy = y0 + r, ix = x0 + tx;
        if (ix < nx && iy < ny) {
            const float a = xf[(size_t)(ix + PAD) + (size_t)NXP * ((size_t)iy + (size_t)ny * pz)];
            img[(size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz)] = a + tile[tx][r];
        }
    }
}

__host__ __device__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct WsPlan {
    size_t trig_off, xf_off, yf_off, total;
    size_t xf_elems, yf_elems;
    int NXP, NYP, NZP;
};

__host__ __device__ WsPlan plan_ws(const int grid[3], int nv)
{
    WsPlan p;
    const size_t nx = grid[0], ny = grid[1], nz = grid[2];
    p.NXP = (int)nx + 2 * PAD;
    p.NYP = (int)ny + 2 * PAD;
    p.NZP = (int)nz + 2 * PAD;
    p.xf_elems = (size_t)p.NXP * ny * (size_t)p.NZP;
    p.yf_elems = (size_t)p.NYP * nx * (size_t)p.NZP;
    p.trig_off = 0;
    p.xf_off = align256(sizeof(double) * 2 * (size_t)(nv > 0 ? nv : 1));
    p.yf_off = p.xf_off + align256(sizeof(float) * p.xf_elems);
    p.total = p.yf_off + align256(sizeof(float) * p.yf_elems);
    return p;
}

bool finite_all(std::initializer_list<double> v)
{
    for (double x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int validate(const int grid[3], const float center[3], const float size[3], const float* angles, int nv,
             float SID, float SAD, int na, float da, int nb, float db, float ai, float bi, unsigned flags)
{
    if (!grid || !center || !size || !angles) return NEFCT_ERR_NULL;
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0 || nv <= 0 || na <= 0 || nb <= 0) return NEFCT_ERR_SHAPE;
    if (nb > 65535 * BY || nv > 65535) return NEFCT_ERR_SHAPE;
    if ((size_t)na * (size_t)nb >= 0x7fffffffull) return NEFCT_ERR_SHAPE;   // 32-bit offsets within a view
    if (grid[0] > 16384 || grid[1] > 16384 || grid[2] > 16384) return NEFCT_ERR_SHAPE;
    const int grid_arr[3] = {grid[0], grid[1], grid[2]};
    const WsPlan p = plan_ws(grid_arr, nv);
    if (p.xf_elems * 4 >= 0x7fffffffull || p.yf_elems * 4 >= 0x7fffffffull) return NEFCT_ERR_SHAPE;
    if (!finite_all({center[0], center[1], center[2], size[0], size[1], size[2], SID, SAD, da, db, ai, bi}))
        return NEFCT_ERR_GEOMETRY;
    if (!(size[0] > 0 && size[1] > 0 && size[2] > 0 && da > 0 && db > 0 && SID > 0)) return NEFCT_ERR_GEOMETRY;
    if (!(SAD > 0)) return NEFCT_ERR_GEOMETRY;
    if (flags & ~(NEFCT_F_INV_L2 | NEFCT_F_REF_SLAB | NEFCT_F_SHORT_SKIP | NEFCT_F_TIE_GT | NEFCT_F_DETERMINISTIC |
                  NEFCT_F_ADJ_WINDOW | NEFCT_F_ADJ_FLOAT_ATOMIC | NEFCT_F_TILE_MASK | NEFCT_F_LANE_STRIDE2 |
                  NEFCT_F_FWD_TILE_MASK | NEFCT_F_FWD_VG_MASK))
        return NEFCT_ERR_UNSUPPORTED;
    if (((flags & NEFCT_F_FWD_VG_MASK) >> NEFCT_F_FWD_VG_SHIFT) > 6) return NEFCT_ERR_UNSUPPORTED;
    if ((flags & NEFCT_F_TILE_MASK) > NEFCT_F_TILE_64x16x16) return NEFCT_ERR_UNSUPPORTED;
    if ((flags & NEFCT_F_DETERMINISTIC) && (flags & NEFCT_F_ADJ_WINDOW)) return NEFCT_ERR_UNSUPPORTED;
    if (flags & NEFCT_F_REF_SLAB) {
        // the reference y-walk treats y slabs as x-voxel thick (siddon.cu.cc:188-201)
        const double vx = (double)size[0] / grid[0], vy = (double)size[1] / grid[1];
        if (fabs(vx - vy) > 1e-6 * vx) return NEFCT_ERR_UNSUPPORTED;
    }
    return NEFCT_OK;
}

Launch make_launch(const int grid[3], const float center[3], const float size[3], int nv, float SID, float SAD,
                   int na, float da, float ai, int nb, float db, float bi, int cyl, unsigned flags, char* ws,
                   const WsPlan& p, const float* zoff = nullptr)
{
    Launch L;
    L.zoff = zoff;
    for (int i = 0; i < 3; ++i) {
        L.V.n[i] = grid[i];
        L.V.vs[i] = (double)size[i] / grid[i];
        L.V.inv_vs[i] = 1.0 / L.V.vs[i];
        L.V.lo[i] = (double)center[i] - (double)size[i] * 0.5;
    }
    L.G.cyl = cyl;
    L.G.sid = SID;
    L.G.sad = SAD;
    L.G.na = na;
    L.G.nb = nb;
    L.G.da = da;
    L.G.ai = ai;
    L.G.db = db;
    L.G.bi = bi;
    L.Ly.xf = (float*)(ws + p.xf_off);
    L.Ly.yf = (float*)(ws + p.yf_off);
    L.Ly.xf_bytes = (uint32_t)(p.xf_elems * 4);
    L.Ly.yf_bytes = (uint32_t)(p.yf_elems * 4);
    L.Ly.NXP = p.NXP;
    L.Ly.NYP = p.NYP;
    L.Ly.NZP = p.NZP;
    L.trig = (const double*)(ws + p.trig_off);
    L.nv = nv;
    L.flags = flags;
    L.sbound = scale_bound(L.V, L.G, flags);
    return L;
}

void launch_trig(const float* angles, int nv, double* trig, hipStream_t st)
{
    hipLaunchKernelGGL(k_view_trig, dim3((nv + 255) / 256), dim3(256), 0, st, angles, nv, trig);
}

void launch_pad(const float* image, const int grid[3], const WsPlan& p, char* ws, hipStream_t st)
{
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    float* xf = (float*)(ws + p.xf_off);
    float* yf = (float*)(ws + p.yf_off);
    hipLaunchKernelGGL(k_pad_x, dim3((p.NXP + 255) / 256, ny, p.NZP), dim3(256), 0, st, image, nx, ny, nz, p.NXP, xf);
    hipLaunchKernelGGL(k_pad_y, dim3((nx + 63) / 64, (p.NYP + 63) / 64, p.NZP), dim3(256), 0, st, image, nx, ny, nz,
                       p.NYP, yf);
}

void launch_zero(float* ptr, size_t n, hipStream_t st)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, ptr, n);
}

int run_forward(const float* image, const int grid[3], const float center[3], const float size[3],
                const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb, float db,
                float bi, int cyl, float* out, unsigned flags, void* workspace, size_t wsbytes, void* stream,
                const float* zoff = nullptr)
{
    int st = validate(grid, center, size, angles, nv, SID, SAD, na, da, nb, db, ai, bi, flags);
    if (st) return st;
    if (!image || !out) return NEFCT_ERR_NULL;
    const WsPlan p = plan_ws(grid, nv);
    if (!workspace || wsbytes < p.total) return NEFCT_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    Launch L = make_launch(grid, center, size, nv, SID, SAD, na, da, ai, nb, db, bi, cyl, flags, ws, p, zoff);
    launch_trig(angles, nv, (double*)(ws + p.trig_off), s);
    launch_pad(image, grid, p, ws, s);
    if (flags & NEFCT_F_REF_SLAB)
        launch_siddon<0, true>(L, nullptr, out, nullptr, s);
    else
        launch_siddon<0, false>(L, nullptr, out, nullptr, s);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

void launch_adjoint_tiles(const Launch& L, const float* pv, float* image, unsigned* maxbits, hipStream_t s);

int run_adjoint(const float* pv, const int grid[3], const float center[3], const float size[3], const float* angles,
                int nv, float SID, float SAD, int na, float da, float ai, int nb, float db, float bi, int cyl,
                float* image, unsigned flags, void* workspace, size_t wsbytes, void* stream,
                const float* zoff = nullptr)
{
    int st = validate(grid, center, size, angles, nv, SID, SAD, na, da, nb, db, ai, bi, flags);
    if (st) return st;
    if (!image || !pv) return NEFCT_ERR_NULL;
    const WsPlan p = plan_ws(grid, nv);
    if (!workspace || wsbytes < p.total) return NEFCT_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    Launch L = make_launch(grid, center, size, nv, SID, SAD, na, da, ai, nb, db, bi, cyl, flags, ws, p, zoff);
    launch_trig(angles, nv, (double*)(ws + p.trig_off), s);
    const int nx = grid[0], ny = grid[1], nz = grid[2];
    const bool ref = (flags & NEFCT_F_REF_SLAB) != 0;
    if (flags & NEFCT_F_ADJ_WINDOW) {
        // ray tiles with LDS slab windows flushed by global atomics
        launch_zero((float*)(ws + p.xf_off), p.xf_elems, s);
        launch_zero((float*)(ws + p.yf_off), p.yf_elems, s);
        const dim3 g((na + BX - 1) / BX, (nb + BY * AR - 1) / (BY * AR), nv), b(BX, BY);
        if (ref)
            hipLaunchKernelGGL((k_siddon_adj<true>), g, b, 0, s, L, pv);
        else
            hipLaunchKernelGGL((k_siddon_adj<false>), g, b, 0, s, L, pv);
        hipLaunchKernelGGL(k_combine, dim3((nx + 63) / 64, (ny + 63) / 64, nz), dim3(256), 0, s,
                           (const float*)(ws + p.xf_off), (const float*)(ws + p.yf_off), nx, ny, nz, p.NXP, p.NYP,
                           image);
    } else {
        // the max |projection| word goes to the otherwise unused padded-copy
        // region of the workspace
        launch_adjoint_tiles(L, pv, image, (unsigned*)(ws + p.xf_off), s);
    }
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

// The volume-tile adjoint (every variant but the slab windows): volume blocks
// own their voxels.  L.trig must hold the view table.  Fixed-point
// accumulation first needs max |projection|, into *maxbits (cleared here).
void launch_adjoint_tiles(const Launch& L, const float* pv, float* image, unsigned* maxbits, hipStream_t s)
{
    const unsigned flags = L.flags;
    const bool ref = (flags & NEFCT_F_REF_SLAB) != 0;
    {
        const bool fix = !(flags & NEFCT_F_ADJ_FLOAT_ATOMIC);
        if (fix) {
            const size_t n = (size_t)L.nv * (size_t)L.G.na * (size_t)L.G.nb;
            launch_zero((float*)maxbits, 1, s);
            const size_t blocks = std::min<size_t>(4096, (n + 255) / 256);
            hipLaunchKernelGGL(k_absmax, dim3((unsigned)blocks), dim3(256), 0, s, pv, n, maxbits);
        }
        const float sb = L.sbound;
        const unsigned shape = flags & NEFCT_F_TILE_MASK;
        const bool cs2 = (flags & NEFCT_F_LANE_STRIDE2) != 0;
        if (!fix && (flags & NEFCT_F_DETERMINISTIC)) {
            // float accumulation in a fixed order: one wavefront per 16x16x8 block
            launch_tile<16, 16, 8, 1, 1>(ref, false, L, pv, image, maxbits, sb, s);
        } else if (shape == NEFCT_F_TILE_32x16x8) {
            if (cs2) launch_tile<32, 16, 8, 4, 2>(ref, fix, L, pv, image, maxbits, sb, s);
            else launch_tile<32, 16, 8, 4, 1>(ref, fix, L, pv, image, maxbits, sb, s);
        } else if (shape == NEFCT_F_TILE_32x32x8) {
            if (cs2) launch_tile<32, 32, 8, 4, 2>(ref, fix, L, pv, image, maxbits, sb, s);
            else launch_tile<32, 32, 8, 4, 1>(ref, fix, L, pv, image, maxbits, sb, s);
        } else if (shape == NEFCT_F_TILE_64x32x8) {
            if (cs2) launch_tile<64, 32, 8, 16, 2>(ref, fix, L, pv, image, maxbits, sb, s);
            else launch_tile<64, 32, 8, 16, 1>(ref, fix, L, pv, image, maxbits, sb, s);
        } else if (shape == NEFCT_F_TILE_16x16x32) {
            if (cs2) launch_tile<16, 16, 32, 4, 2>(ref, fix, L, pv, image, maxbits, sb, s);
            else launch_tile<16, 16, 32, 4, 1>(ref, fix, L, pv, image, maxbits, sb, s);
        } else if (shape == NEFCT_F_TILE_64x16x16) {
            if (cs2) launch_tile<64, 16, 16, 16, 2>(ref, fix, L, pv, image, maxbits, sb, s);
            else launch_tile<64, 16, 16, 16, 1>(ref, fix, L, pv, image, maxbits, sb, s);
        } else {
            if (cs2) launch_tile<32, 32, 16, 16, 2>(ref, fix, L, pv, image, maxbits, sb, s);
            else launch_tile<32, 32, 16, 16, 1>(ref, fix, L, pv, image, maxbits, sb, s);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched problems (one launch per direction over the tasks of every problem)
// ---------------------------------------------------------------------------
constexpr int BTX = 32, BTY = 32, BTZ = 16, BNW = 16;  // adjoint block: the plain default
constexpr int Q_HEADER = 256;            // workspace head (reserved), then one Launch per problem

__device__ Launch launch_of(const nefct_problem& pr, char* ws_base, unsigned flags)
{
    Launch L;
    L.zoff = nullptr;
    for (int i = 0; i < 3; ++i) {
        L.V.n[i] = pr.grid[i];
        L.V.vs[i] = (double)pr.size[i] / pr.grid[i];
        L.V.inv_vs[i] = 1.0 / L.V.vs[i];
        L.V.lo[i] = (double)pr.center[i] - (double)pr.size[i] * 0.5;
    }
    L.G.cyl = pr.cyl;
    L.G.sid = pr.SID;
    L.G.sad = pr.SAD;
    L.G.na = pr.na;
    L.G.nb = pr.nb;
    L.G.da = pr.da;
    L.G.ai = pr.ai;
    L.G.db = pr.db;
    L.G.bi = pr.bi;
    const WsPlan p = plan_ws(pr.grid, pr.nv);
    char* ws = ws_base + pr.ws_offset;
    L.Ly.xf = (float*)(ws + p.xf_off);
    L.Ly.yf = (float*)(ws + p.yf_off);
    L.Ly.xf_bytes = (uint32_t)(p.xf_elems * 4);
    L.Ly.yf_bytes = (uint32_t)(p.yf_elems * 4);
    L.Ly.NXP = p.NXP;
    L.Ly.NYP = p.NYP;
    L.Ly.NZP = p.NZP;
    L.trig = (const double*)(ws + p.trig_off);
    L.nv = pr.nv;
    L.flags = flags;
    L.sbound = scale_bound(L.V, L.G, flags);
    return L;
}

__device__ __forceinline__ int find_problem(const nefct_problem* __restrict__ probs, int nprob, long long t,
                                            bool blocks)
{
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long long b = blocks ? probs[mid].block_begin : probs[mid].tile_begin;
        if (b <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__host__ __device__ inline size_t batch_launch_bytes(int nprob) { return align256(sizeof(Launch) * (size_t)nprob); }

// Per-problem launch parameters live in the workspace: a kernel that picks its
// problem at run time reads them from global memory (uniform scalar loads)
// instead of building a private copy, which would be indexed dynamically and
// so be placed in scratch memory.  After them, one max |projection| word per
// problem (fixed-point adjoint scale).
__device__ __forceinline__ const Launch* batch_launches(char* ws_base) { return (const Launch*)(ws_base + Q_HEADER); }
__host__ __device__ inline size_t batch_max_off(int nprob) { return Q_HEADER + batch_launch_bytes(nprob); }
__host__ __device__ inline size_t batch_head_bytes(int nprob) { return batch_max_off(nprob) + align256(4 * (size_t)nprob); }

// trig tables and launch parameters of every problem; zeroed max words
__global__ void k_batch_prep(const nefct_problem* __restrict__ probs, int nprob, char* ws_base, unsigned flags)
{
    const int p = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob) return;
    const nefct_problem& pr = probs[p];
    if (i == 0) {
        ((Launch*)(ws_base + Q_HEADER))[p] = launch_of(pr, ws_base, flags);
        ((unsigned*)(ws_base + batch_max_off(nprob)))[p] = 0u;
    }
    if (i >= pr.nv) return;
    const WsPlan pl = plan_ws(pr.grid, pr.nv);
    double* trig = (double*)(ws_base + pr.ws_offset + pl.trig_off);
    const double a = (double)pr.angles[i];
    trig[2 * i] = cos(a);
    trig[2 * i + 1] = sin(a);
}

// Forward of every problem in one launch.  Each problem's tasks are numbered
// as in k_siddon (view fastest within groups of FWD_VG views, then ray tile,
// then view group; the last group padded to FWD_VG), problems one after the
// other from tile_begin, so every problem's task count is a multiple of 8.
// Workgroup ids follow the same order problem by problem, and within a
// problem they are remapped so that each of the 8 XCDs gets a contiguous
// eighth of its tasks: all XCDs work on the same problem at a time (one
// volume in the shared cache, as in one plain call), each on its own views.
// t0 (a multiple of 8) offsets the chunk this launch covers.
template <bool REF>
__global__ __launch_bounds__(FX * FY) void k_batch_fwd(const nefct_problem* __restrict__ probs, int nprob,
                                                       long long t0, long long total, char* ws_base)
{
    const long long id = t0 + blockIdx.x;
    if (id >= total) return;
    const int p = __builtin_amdgcn_readfirstlane(find_problem(probs, nprob, id, false));
    const nefct_problem& pr = probs[p];
    const long long begin = pr.tile_begin;
    const long long n = (p + 1 < nprob ? probs[p + 1].tile_begin : total) - begin;
    const long long lid = id - begin;
    const long long local = (lid & 7) * (n >> 3) + (lid >> 3);
    const unsigned tiles = (unsigned)pr.tiles_a * (unsigned)pr.tiles_b;
    const unsigned vin = (unsigned)(local % FWD_VG);
    const long long rest = local / FWD_VG;
    const unsigned tile = (unsigned)(rest % tiles);
    const int iv = (int)((rest / tiles) * FWD_VG + vin);
    if (iv >= pr.nv) return;
    const int ta = (int)(tile % (unsigned)pr.tiles_a), tb = (int)(tile / (unsigned)pr.tiles_a);
    const Launch& L = batch_launches(ws_base)[p];
    siddon_rays<0, REF>(L, ta * FX + threadIdx.x, tb * FY + threadIdx.y, iv, nullptr, pr.sinogram, nullptr);
}

// max |sinogram| of every problem (blockIdx.y = problem)
__global__ __launch_bounds__(256) void k_batch_absmax(const nefct_problem* __restrict__ probs, int nprob,
                                                      char* ws_base)
{
    const int p = blockIdx.y;
    if (p >= nprob) return;
    const nefct_problem& pr = probs[p];
    const size_t n = (size_t)pr.nv * (size_t)pr.na * (size_t)pr.nb;
    const float* x = pr.sinogram;
    unsigned m = 0u;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        m = max(m, __float_as_uint(x[i]) & 0x7fffffffu);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax((unsigned*)(ws_base + batch_max_off(nprob)) + p, m);
}

// One launch over the voxel blocks of every problem: workgroup g takes global
// block g and finds its problem by binary search over block_begin.  (A
// persistent loop over an atomic queue holds enough extra state across
// iterations to make this kernel spill at 16 wavefronts; the blocks are equal
// in size within a problem, so the dispatcher balances them as well.)
template <bool REF, int CS>
__global__ __launch_bounds__(64 * BNW) void k_batch_adj(const nefct_problem* __restrict__ probs, int nprob,
                                                        long long total, char* ws_base, unsigned flags)
{
    __shared__ BlockAcc<BTX, BTY, BTZ, true> acc;
    __shared__ int4 shadows[SHV];
    const long long t = (long long)blockIdx.x + 65536ll * (long long)blockIdx.y;
    if (t >= total) return;
    const int p = __builtin_amdgcn_readfirstlane(find_problem(probs, nprob, t, true));
    const nefct_problem& pr = probs[p];
    // block coordinates made scalar: in vector registers they would carry the
    // block bounds and their derived values into VGPRs
    const int local = __builtin_amdgcn_readfirstlane((int)(t - pr.block_begin));
    const int bx = __builtin_amdgcn_readfirstlane(local % pr.blocks[0]);
    const int by = __builtin_amdgcn_readfirstlane((local / pr.blocks[0]) % pr.blocks[1]);
    const int bz = __builtin_amdgcn_readfirstlane(local / (pr.blocks[0] * pr.blocks[1]));
    const Launch& L = batch_launches(ws_base)[p];
    const int es = fixed_exp(((const unsigned*)(ws_base + batch_max_off(nprob)))[p], L.sbound);
    adj_tile_block<REF, true, BTX, BTY, BTZ, BNW, CS>(L, pr.sinogram, const_cast<float*>(pr.volume), bx, by, bz, acc,
                                                     shadows, es, false);
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free adjoint: its scratch in a fixed device array
// ---------------------------------------------------------------------------
// The volume-tile adjoint needs, beyond its inputs and output, only the
// max |projection| word of the fixed-point scale and the per-view (cos, sin)
// table.  For callers that cannot allocate (graph capture, serving loops)
// both live in `scratch`: bytes [0, 256) hold the max word, then 16 bytes per
// view for at most NOALLOC_MAX_VIEWS views (128 KB).  Every device holds its
// own copy until the process ends; in host code `scratch` names a host-side
// stand-in, so its device address is resolved per device with
// hipGetSymbolAddress and only that address is passed to kernels.  Calls that
// overlap on one device share the array and race.
constexpr int NOALLOC_MAX_VIEWS = 8192;
constexpr size_t NOALLOC_TRIG_OFF = 256;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_TRIG_OFF + 2 * sizeof(double) * (size_t)NOALLOC_MAX_VIEWS;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch_dev[NOALLOC_MAX_DEVICES];   // device address of `scratch`, per device

// Device address of `scratch` on the current device, resolved on the first
// call made with that device current.
int scratch_of_current_device(char** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return NEFCT_ERR_LAUNCH;
    if (dev >= NOALLOC_MAX_DEVICES) return NEFCT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NEFCT_ERR_LAUNCH;
        g_scratch_dev[dev] = (char*)p;
    }
    *out = g_scratch_dev[dev];
    return NEFCT_OK;
}

int run_adjoint_noalloc(const float* pv, const int grid[3], const float center[3], const float size[3],
                        const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb,
                        float db, float bi, int cyl, float* image, unsigned flags, void* stream)
{
    int st = validate(grid, center, size, angles, nv, SID, SAD, na, da, nb, db, ai, bi, flags);
    if (st) return st;
    if (!image || !pv) return NEFCT_ERR_NULL;
    if (nv > NOALLOC_MAX_VIEWS) return NEFCT_ERR_SCRATCH;
    if (flags & NEFCT_F_ADJ_WINDOW) return NEFCT_ERR_UNSUPPORTED;   // needs the padded-copy workspace
    char* sc = nullptr;
    st = scratch_of_current_device(&sc);
    if (st) return st;
    // the tile kernels read the view table and the max word only; the padded
    // copies of the plan are never touched (their strides still describe the
    // walk axes)
    WsPlan p = plan_ws(grid, nv);
    p.trig_off = NOALLOC_TRIG_OFF;
    p.xf_off = p.yf_off = 0;
    p.xf_elems = p.yf_elems = 0;
    Launch L = make_launch(grid, center, size, nv, SID, SAD, na, da, ai, nb, db, bi, cyl, flags, sc, p);
    hipStream_t s = (hipStream_t)stream;
    launch_trig(angles, nv, (double*)(sc + NOALLOC_TRIG_OFF), s);
    launch_adjoint_tiles(L, pv, image, (unsigned*)sc, s);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // namespace

extern "C" {

const char* nefct_status_string(int status)
{
    switch (status) {
    case NEFCT_OK: return "ok";
    case NEFCT_ERR_NULL: return "null pointer argument";
    case NEFCT_ERR_SHAPE: return "invalid or oversized shape";
    case NEFCT_ERR_GEOMETRY: return "non-finite or degenerate geometry";
    case NEFCT_ERR_WORKSPACE: return "workspace missing or too small";
    case NEFCT_ERR_UNSUPPORTED: return "unsupported flags or configuration";
    case NEFCT_ERR_LAUNCH: return "HIP kernel launch failed";
    case NEFCT_ERR_SCRATCH: return "beyond the fixed scratch of the allocation-free entry point (at most 8192 views)";
    default: return "unknown status";
    }
}

size_t nefct_siddon_workspace_size(const int grid[3], int nv)
{
    if (!grid || grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0) return 0;
    return plan_ws(grid, nv).total;
}

int nefct_project_flat(const float* image, const int grid[3], const float center[3], const float size[3],
                       const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb,
                       float db, float bi, float* projection_value, unsigned flags, void* workspace,
                       size_t workspace_bytes, void* stream)
{
    return run_forward(image, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, 0,
                       projection_value, flags, workspace, workspace_bytes, stream);
}

int nefct_project_cyli(const float* image, const int grid[3], const float center[3], const float size[3],
                       const float* angles, int nv, float SID, float SAD, int na, float da, float ai, int nb,
                       float db, float bi, float* projection_value, unsigned flags, void* workspace,
                       size_t workspace_bytes, void* stream)
{
    return run_forward(image, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, 1,
                       projection_value, flags, workspace, workspace_bytes, stream);
}

int nefct_backproject_flat(const float* projection_value, const int grid[3], const float center[3],
                           const float size[3], const float* angles, int nv, float SID, float SAD, int na, float da,
                           float ai, int nb, float db, float bi, float* image, unsigned flags, void* workspace,
                           size_t workspace_bytes, void* stream)
{
    return run_adjoint(projection_value, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, 0, image,
                       flags, workspace, workspace_bytes, stream);
}

int nefct_backproject_flat_noalloc(const float* projection_value, const int grid[3], const float center[3],
                                   const float size[3], const float* angles, int nv, float SID, float SAD, int na,
                                   float da, float ai, int nb, float db, float bi, float* image, unsigned flags,
                                   void* stream)
{
    return run_adjoint_noalloc(projection_value, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, 0,
                               image, flags, stream);
}

int nefct_backproject_cyli(const float* projection_value, const int grid[3], const float center[3],
                           const float size[3], const float* angles, int nv, float SID, float SAD, int na, float da,
                           float ai, int nb, float db, float bi, float* image, unsigned flags, void* workspace,
                           size_t workspace_bytes, void* stream)
{
    return run_adjoint(projection_value, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, 1, image,
                       flags, workspace, workspace_bytes, stream);
}

int nefct_project_helical(const float* image, const int grid[3], const float center[3], const float size[3],
                          const float* angles, const float* offsets, int nv, float SID, float SAD, int na, float da,
                          float ai, int nb, float db, float bi, int cyl, float* projection_value, unsigned flags,
                          void* workspace, size_t workspace_bytes, void* stream)
{
    if (!offsets) return NEFCT_ERR_NULL;
    return run_forward(image, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi, cyl ? 1 : 0,
                       projection_value, flags, workspace, workspace_bytes, stream, offsets);
}

int nefct_backproject_helical(const float* projection_value, const int grid[3], const float center[3],
                              const float size[3], const float* angles, const float* offsets, int nv, float SID,
                              float SAD, int na, float da, float ai, int nb, float db, float bi, int cyl,
                              float* image, unsigned flags, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!offsets) return NEFCT_ERR_NULL;
    return run_adjoint(projection_value, grid, center, size, angles, nv, SID, SAD, na, da, ai, nb, db, bi,
                       cyl ? 1 : 0, image, flags, workspace, workspace_bytes, stream, offsets);
}

int nefct_siddon_count(const int grid[3], const float center[3], const float size[3], const float* angles, int nv,
                       float SID, float SAD, int na, float da, float ai, int nb, float db, float bi, int cyl,
                       int32_t* counts, unsigned flags, void* workspace, size_t workspace_bytes, void* stream)
{
    int st = validate(grid, center, size, angles, nv, SID, SAD, na, da, nb, db, ai, bi, flags);
    if (st) return st;
    if (!counts) return NEFCT_ERR_NULL;
    const WsPlan p = plan_ws(grid, nv);
    if (!workspace || workspace_bytes < p.total) return NEFCT_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    Launch L = make_launch(grid, center, size, nv, SID, SAD, na, da, ai, nb, db, bi, cyl, flags, ws, p);
    launch_trig(angles, nv, (double*)(ws + p.trig_off), s);
    if (flags & NEFCT_F_REF_SLAB)
        launch_siddon<2, true>(L, nullptr, nullptr, counts, s);
    else
        launch_siddon<2, false>(L, nullptr, nullptr, counts, s);
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_batch_plan(nefct_problem* probs, int nprob, unsigned flags, size_t* workspace_bytes, long long* total_tiles,
                     long long* total_blocks)
{
    if (!probs || !workspace_bytes || !total_tiles || !total_blocks) return NEFCT_ERR_NULL;
    if (nprob <= 0 || nprob > 65535) return NEFCT_ERR_SHAPE;
    long long tiles = 0, blocks = 0;
    size_t off = batch_head_bytes(nprob);
    for (int i = 0; i < nprob; ++i) {
        nefct_problem& pr = probs[i];
        int st = validate(pr.grid, pr.center, pr.size, pr.angles, pr.nv, pr.SID, pr.SAD, pr.na, pr.da, pr.nb, pr.db,
                          pr.ai, pr.bi, flags);
        if (st) return st;
        if (!pr.volume || !pr.sinogram) return NEFCT_ERR_NULL;
        pr.tiles_a = (pr.na + FX - 1) / FX;
        pr.tiles_b = (pr.nb + FY - 1) / FY;
        pr.tile_begin = tiles;
        tiles += (long long)pr.tiles_a * pr.tiles_b * ((pr.nv + FWD_VG - 1) / FWD_VG * FWD_VG);
        pr.blocks[0] = (pr.grid[0] + BTX - 1) / BTX;
        pr.blocks[1] = (pr.grid[1] + BTY - 1) / BTY;
        pr.blocks[2] = (pr.grid[2] + BTZ - 1) / BTZ;
        pr.block_begin = blocks;
        blocks += (long long)pr.blocks[0] * pr.blocks[1] * pr.blocks[2];
        pr.ws_offset = (long long)off;
        off += align256(plan_ws(pr.grid, pr.nv).total);
        pr.pad_ = 0;
    }
    *workspace_bytes = off;
    *total_tiles = tiles;
    *total_blocks = blocks;
    return NEFCT_OK;
}

static int batch_prep(const nefct_problem* host, const nefct_problem* dev, int nprob, char* ws, unsigned flags,
                      hipStream_t s)
{
    int maxnv = 1;
    for (int i = 0; i < nprob; ++i) maxnv = max(maxnv, host[i].nv);
    hipLaunchKernelGGL(k_batch_prep, dim3((maxnv + 255) / 256, nprob), dim3(256), 0, s, dev, nprob, ws, flags);
    return NEFCT_OK;
}

int nefct_batch_project(const nefct_problem* host, const nefct_problem* dev, int nprob, long long total_tiles,
                        unsigned flags, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!host || !dev || !workspace) return NEFCT_ERR_NULL;
    if (nprob <= 0) return NEFCT_ERR_SHAPE;
    const nefct_problem& last = host[nprob - 1];
    if (workspace_bytes < (size_t)last.ws_offset + align256(plan_ws(last.grid, last.nv).total))
        return NEFCT_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    batch_prep(host, dev, nprob, ws, flags, s);
    for (int i = 0; i < nprob; ++i) {
        const WsPlan p = plan_ws(host[i].grid, host[i].nv);
        launch_pad(host[i].volume, host[i].grid, p, ws + host[i].ws_offset, s);
    }
    // chunks of at most 2^31 / 256 workgroups (1-D grid of 32-bit work-items)
    const long long chunk = (0x7fffffffll / (FX * FY)) / 8 * 8;
    for (long long t0 = 0; t0 < total_tiles; t0 += chunk) {
        const long long blocks = std::min(chunk, total_tiles - t0);
        if (flags & NEFCT_F_REF_SLAB)
            hipLaunchKernelGGL((k_batch_fwd<true>), dim3((unsigned)blocks), dim3(FX, FY), 0, s, dev, nprob, t0,
                               total_tiles, ws);
        else
            hipLaunchKernelGGL((k_batch_fwd<false>), dim3((unsigned)blocks), dim3(FX, FY), 0, s, dev, nprob, t0,
                               total_tiles, ws);
    }
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

int nefct_batch_backproject(const nefct_problem* host, const nefct_problem* dev, int nprob, long long total_blocks,
                            unsigned flags, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!host || !dev || !workspace) return NEFCT_ERR_NULL;
    if (nprob <= 0) return NEFCT_ERR_SHAPE;
    const nefct_problem& last = host[nprob - 1];
    if (workspace_bytes < (size_t)last.ws_offset + align256(plan_ws(last.grid, last.nv).total))
        return NEFCT_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    batch_prep(host, dev, nprob, ws, flags, s);
    size_t maxn = 1;
    for (int i = 0; i < nprob; ++i)
        maxn = std::max(maxn, (size_t)host[i].nv * (size_t)host[i].na * (size_t)host[i].nb);
    const unsigned bx = (unsigned)std::min<size_t>(1024, (maxn + 255) / 256);
    hipLaunchKernelGGL(k_batch_absmax, dim3(bx, nprob), dim3(256), 0, s, dev, nprob, ws);
    if (total_blocks <= 0) return NEFCT_OK;
    if (total_blocks > 65536ll * 65535ll) return NEFCT_ERR_SHAPE;
    const dim3 g((unsigned)std::min<long long>(total_blocks, 65536), (unsigned)((total_blocks + 65535) / 65536));
    // lanes one or two detector rows apart (NEFCT_F_LANE_STRIDE2), as in the plain adjoint
    const bool cs2 = (flags & NEFCT_F_LANE_STRIDE2) != 0;
    if (flags & NEFCT_F_REF_SLAB) {
        if (cs2)
            hipLaunchKernelGGL((k_batch_adj<true, 2>), g, dim3(64 * BNW), 0, s, dev, nprob, total_blocks, ws, flags);
        else
            hipLaunchKernelGGL((k_batch_adj<true, 1>), g, dim3(64 * BNW), 0, s, dev, nprob, total_blocks, ws, flags);
    } else {
        if (cs2)
            hipLaunchKernelGGL((k_batch_adj<false, 2>), g, dim3(64 * BNW), 0, s, dev, nprob, total_blocks, ws, flags);
        else
            hipLaunchKernelGGL((k_batch_adj<false, 1>), g, dim3(64 * BNW), 0, s, dev, nprob, total_blocks, ws, flags);
    }
    NEFCT_CHECK_LAUNCH();
    return NEFCT_OK;
}

}  // extern "C"
