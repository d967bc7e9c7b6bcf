"""Known answers that pin the oracle's reference-rule restatement and the FDK
backprojection's orientation independently of the restatements themselves.

* Q2: the reference's attribution of a slab with a mixed-sign y/z (or x/z)
  double crossing, in both branches of the x-slab walk
  (nefct/cc/siddon_3d/siddon.cu.cc:132-176) and of the y-slab walk (:272-316),
  compared voxel by voxel with weights worked out by hand (tests/_kat_siddon.py).
  Swapping any one of the three pieces of either branch changes a voxel.
* Q3 (:472): the short-ray skip just below and above 10 voxels.
* Q6 (:41 vs :480): an exactly tied |xd| = |yd| ray walks x slabs in the
  forward and y slabs in the adjoint.
* FDK: the voxel-to-detector map of ``fdk_backproject`` against the line
  through the source and the voxel (built from ``ray_points``, the exact
  Siddon oracle's own ray set-up), and an off-centre ball reconstructed at its
  own place, not at any mirror image.
"""
import numpy as np
import pytest

import _kat_siddon as K


def _one_ray(oracle, case, SID=K.SID, SAD=K.SAD):
    V = oracle.vol_struct(case["grid"], case["center"], case["grid"])   # 1 mm voxels
    na, da, ai, nb, db, bi = K.det_args(case["a"], case["b"])
    det = oracle.det_struct(0, SID, SAD, na, da, ai, nb, db, bi)
    return V, det


def _unit_sino():
    return np.ones((1, 1, 1))


@pytest.mark.parametrize("name", sorted(K.Q2_CASES))
def test_q2_hand_derived_attribution(oracle, name):
    case = K.Q2_CASES[name]
    V, det = _one_ray(oracle, case)
    ang = np.zeros(1)
    want_ref = K.dense(case["grid"], case["ref"])
    want_exact = K.dense(case["grid"], case["exact"])
    assert not np.allclose(want_ref, want_exact)
    # reference rule: FP64 and FP32 instantiations of the restatement
    got64 = oracle.ref_bp64(V, det, ang, _unit_sino(), flags=oracle.REF_BP_FLAGS)
    assert np.abs(got64 - want_ref).max() <= 1e-12
    got32 = oracle.ref_bp(V, det, ang.astype(np.float32), _unit_sino(), flags=oracle.REF_BP_FLAGS)
    assert np.abs(got32 - want_ref).max() <= 2e-6
    # the same restatement with the attribution corrected, and the exact oracle
    assert np.abs(oracle.ref_bp64(V, det, ang, _unit_sino(), flags=0) - want_exact).max() <= 1e-12
    assert np.abs(oracle.exact_bp(V, det, ang, _unit_sino()) - want_exact).max() <= 1e-12
    # the reference forward reads the same weights, divided by L^2 (Q1)
    nx, ny, nz = case["grid"]
    for (z, y, x), w in case["ref"].items():
        img = np.zeros((nz, ny, nx))
        img[z, y, x] = 1.0
        f = oracle.ref_fwd64(V, det, ang, img, flags=oracle.REF_FWD_FLAGS)[0, 0, 0]
        assert f == pytest.approx(w / K.Q2_L2[name], rel=1e-12)


def test_q3_short_ray_skip(oracle):
    c = K.Q3_CASE
    ang = np.zeros(1)
    want = K.dense(c["grid"], c["hit"])
    for SAD, kept in ((K.Q3_SAD_SKIPPED, False), (K.Q3_SAD_KEPT, True)):
        V, det = _one_ray(oracle, c, SID=c["SID"], SAD=SAD)
        got = oracle.ref_bp64(V, det, ang, _unit_sino(), flags=oracle.REF_BP_FLAGS)
        assert np.abs(got - (want if kept else 0.0)).max() <= 1e-12, SAD
        got32 = oracle.ref_bp(V, det, ang.astype(np.float32), _unit_sino(), flags=oracle.REF_BP_FLAGS)
        assert np.abs(got32 - (want if kept else 0.0)).max() <= 1e-6, SAD
        # without the toggle (and in the forward) the ray always counts
        assert np.abs(oracle.ref_bp64(V, det, ang, _unit_sino(), flags=oracle.REF_SLAB) - want).max() <= 1e-12
        f = oracle.ref_fwd64(V, det, ang, np.ones(want.shape), flags=oracle.REF_FWD_FLAGS)[0, 0, 0]
        assert f == pytest.approx(4.0 / SAD ** 2, rel=1e-12)


def test_q6_tied_ray_walks_differ(oracle):
    c = K.Q6_CASE
    V, det = _one_ray(oracle, c)
    ang = np.zeros(1)
    fwd_w = K.dense(c["grid"], c["fwd_x_walk"])
    adj_w = K.dense(c["grid"], c["adj_y_walk"])
    # adjoint with the reference's '>' rule: the y walk
    got = oracle.ref_bp64(V, det, ang, _unit_sino(), flags=oracle.REF_BP_FLAGS)
    assert np.abs(got - adj_w).max() <= 1e-12
    got32 = oracle.ref_bp(V, det, ang.astype(np.float32), _unit_sino(), flags=oracle.REF_BP_FLAGS)
    assert np.abs(got32 - adj_w).max() <= 2e-6
    # '>=' (no TIE_GT): the x walk, the forward's
    got = oracle.ref_bp64(V, det, ang, _unit_sino(), flags=oracle.REF_SLAB | oracle.SHORT_SKIP)
    assert np.abs(got - fwd_w).max() <= 1e-12
    for z in range(4):
        img = np.zeros(fwd_w.shape)
        img[z, 0, 0] = 1.0
        f = oracle.ref_fwd64(V, det, ang, img, flags=oracle.REF_FWD_FLAGS)[0, 0, 0]
        assert f == pytest.approx(fwd_w[z, 0, 0] / K.Q6_L2, rel=1e-12, abs=1e-18)
    # the exact transform does not depend on the walk axis
    assert np.abs(oracle.exact_bp(V, det, ang, _unit_sino()) - K.dense(c["grid"], c["exact"])).max() <= 1e-12


def _line_hit(oracle, det, t, P):
    """Fractional detector indices (fa, fb) where the line from the source
    through P meets the detector plane, and lambda = |hit - S| / |P - S|, from
    the Siddon oracle's ray end points of cells (0,0), (1,0), (0,1)."""
    S, D00 = oracle.ray_points(det, t, 0, 0)
    _, D10 = oracle.ray_points(det, t, 1, 0)
    _, D01 = oracle.ray_points(det, t, 0, 1)
    # S + lam (P - S) = D00 + fa (D10 - D00) + fb (D01 - D00)
    M = np.stack([P - S, -(D10 - D00), -(D01 - D00)], axis=1)
    lam, fa, fb = np.linalg.solve(M, D00 - S)
    return fa, fb, lam


def test_fdk_backprojection_maps_voxels_along_source_rays(oracle):
    """A sinogram linear in the detector indices is reproduced exactly by
    bilinear sampling, so one view's backprojection at voxel P must equal
    (SOD/U)^2 (alpha fa + beta fb + gamma) with (fa, fb) where the source ray
    through P meets the detector and U = SDD / lambda.  A mirrored a or b
    axis, a wrong magnification or a half-cell shift fails it."""
    vol = oracle.vol_struct((7, 6, 5), (3.0, -2.0, 1.5), (7 * 4.0, 6 * 4.0, 5 * 3.0))
    det = oracle.det_struct(0, 300.0, 480.0, 41, 2.0, 0.3, 37, 1.8, 0.7)
    alpha, beta, gamma = 0.37, -0.21, 5.0
    ia = np.arange(det.na)
    ib = np.arange(det.nb)
    q = (alpha * ia[None, :] + beta * ib[:, None] + gamma)[None]
    x = vol.cx - vol.sx / 2 + (np.arange(vol.nx) + 0.5) * vol.sx / vol.nx
    y = vol.cy - vol.sy / 2 + (np.arange(vol.ny) + 0.5) * vol.sy / vol.ny
    z = vol.cz - vol.sz / 2 + (np.arange(vol.nz) + 0.5) * vol.sz / vol.nz
    for t in (0.0, 0.7, 2.4, 4.1):
        bp = oracle.fdk_backproject(vol, det, np.array([t]), q)
        checked = 0
        for iz in range(vol.nz):
            for iy in range(vol.ny):
                for ix in range(vol.nx):
                    P = np.array([x[ix], y[iy], z[iz]])
                    fa, fb, lam = _line_hit(oracle, det, t, P)
                    if not (0 <= fa <= det.na - 1 and 0 <= fb <= det.nb - 1):
                        continue
                    U = det.sad / lam
                    want = (det.sid / U) ** 2 * (alpha * fa + beta * fb + gamma)
                    assert bp[iz, iy, ix] == pytest.approx(want, rel=1e-10), (t, ix, iy, iz)
                    checked += 1
        assert checked > 100


def _ball(vol_shape, vs, centre, radius, mu):
    n = vol_shape
    g = [(np.arange(n[i]) + 0.5) * vs - n[i] * vs / 2 for i in range(3)]
    Z, Y, X = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    return mu * (((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) <= radius ** 2)


def ball_region_means(rec, n, vs, c, half=2):
    """Mean of rec in a (2 half)^3 cube around point c (mm, volume centred)."""
    i = [int(round(c[k] / vs + n / 2)) for k in range(3)]
    return rec[i[2] - half:i[2] + half, i[1] - half:i[1] + half, i[0] - half:i[0] + half].mean()


def test_fdk_off_centre_ball_not_mirrored(oracle):
    """FDK of the exact projections of a ball centred at (14, -10, 8) mm
    returns its attenuation there and ~0 at each mirror image (x, y or z
    negated, and x <-> y swapped)."""
    n, vs = 32, 2.0
    V = oracle.vol_struct((n, n, n), (0, 0, 0), (n * vs,) * 3)
    c = (14.0, -10.0, 8.0)
    mu = 0.02
    img = _ball((n, n, n), vs, c, 12.0, mu)
    det = oracle.det_struct(0, 300.0, 450.0, 96, 1.5, 0.5, 96, 1.5, 0.5)
    nv = 120
    ang = 2 * np.pi * np.arange(nv) / nv
    sino = oracle.exact_fwd(V, det, ang, img)
    q = sino * oracle.cosine_weight(det)[None]
    q = oracle.ramp_filter_rows(q) * oracle.fdk_scale(det, nv, 2 * np.pi / nv, False)
    rec = oracle.fdk_backproject(V, det, ang, q)
    assert ball_region_means(rec, n, vs, c) == pytest.approx(mu, rel=0.05)
    for m in ((-c[0], c[1], c[2]), (c[0], -c[1], c[2]), (c[0], c[1], -c[2]), (c[1], c[0], c[2])):
        assert abs(ball_region_means(rec, n, vs, m)) < 0.1 * mu, m
