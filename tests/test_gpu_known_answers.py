"""The GPU kernels against hand-derived known answers (no restatement in the
loop): reference-mode slab attribution in both branches of both walks (Q2),
the short-ray skip (Q3) and the tie rule (Q6) of the reference adjoint, the
exact mode on the same rays, and FDK orientation with an off-centre ball.
The weights are listed, with their derivation, in tests/_kat_siddon.py.
"""
import math

import numpy as np
import pytest

import _kat_siddon as K

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _geom(case, SID=K.SID, SAD=K.SAD):
    import nefct_hip as nh
    na, da, ai, nb, db, bi = K.det_args(case["a"], case["b"])
    vol = nh.VolumeGeometry(tuple(case["grid"]), tuple(case["center"]), tuple(float(g) for g in case["grid"]))
    det = nh.ConeGeometry("flat", SID, SAD, na, da, ai, nb, db, bi)
    return vol, det


def _bp(case, flags, SID=K.SID, SAD=K.SAD):
    import nefct_hip as nh
    vol, det = _geom(case, SID, SAD)
    y = torch.ones((1, 1, 1), device=DEV)
    return nh.backproject(y, vol, det, torch.zeros(1), flags).double().cpu().numpy()


def _fwd_weights(case, flags, SID=K.SID, SAD=K.SAD):
    """Forward of each one-hot volume: the ray's weight per voxel."""
    import nefct_hip as nh
    vol, det = _geom(case, SID, SAD)
    nx, ny, nz = case["grid"]
    eye = torch.eye(nx * ny * nz, device=DEV).reshape(nx * ny * nz, nz, ny, nx)
    out = np.zeros((nz, ny, nx))
    for k in range(nx * ny * nz):
        out.reshape(-1)[k] = nh.project(eye[k], vol, det, torch.zeros(1), flags)[0, 0, 0].item()
    return out


@pytest.mark.parametrize("name", sorted(K.Q2_CASES))
def test_q2_reference_mode_hand_derived(name):
    from nefct_hip._lib import REF_ADJOINT, REF_FORWARD
    case = K.Q2_CASES[name]
    ref = K.dense(case["grid"], case["ref"])
    exact = K.dense(case["grid"], case["exact"])
    assert np.abs(_bp(case, REF_ADJOINT) - ref).max() <= 2e-6
    assert np.abs(_bp(case, 0) - exact).max() <= 2e-6
    assert np.abs(_fwd_weights(case, REF_FORWARD) * K.Q2_L2[name] - ref).max() <= 2e-6
    assert np.abs(_fwd_weights(case, 0) - exact).max() <= 2e-6


def test_q3_reference_adjoint_short_ray_skip():
    from nefct_hip._lib import F_REF_SLAB, REF_ADJOINT
    c = K.Q3_CASE
    hit = K.dense(c["grid"], c["hit"])
    assert np.abs(_bp(c, REF_ADJOINT, c["SID"], K.Q3_SAD_SKIPPED)).max() == 0.0
    assert np.abs(_bp(c, REF_ADJOINT, c["SID"], K.Q3_SAD_KEPT) - hit).max() <= 1e-6
    assert np.abs(_bp(c, F_REF_SLAB, c["SID"], K.Q3_SAD_SKIPPED) - hit).max() <= 1e-6


def test_q6_reference_tie_rule():
    from nefct_hip._lib import F_REF_SLAB, F_SHORT_SKIP, REF_ADJOINT, REF_FORWARD
    c = K.Q6_CASE
    assert np.abs(_bp(c, REF_ADJOINT) - K.dense(c["grid"], c["adj_y_walk"])).max() <= 2e-6
    assert np.abs(_bp(c, F_REF_SLAB | F_SHORT_SKIP) - K.dense(c["grid"], c["fwd_x_walk"])).max() <= 2e-6
    fw = _fwd_weights(c, REF_FORWARD) * K.Q6_L2
    assert np.abs(fw - K.dense(c["grid"], c["fwd_x_walk"])).max() <= 2e-6
    assert np.abs(_bp(c, 0) - K.dense(c["grid"], c["exact"])).max() <= 2e-6


def test_fdk_off_centre_ball_not_mirrored_gpu(oracle):
    """GPU forward + GPU FDK of a ball centred at (30, -20, 16) mm: its value
    there, ~0 at the mirror images, and the backprojection equal to the
    oracle's on the same filtered projections."""
    import nefct_hip as nh
    from nefct_hip import fdk as F
    n, vs = 64, 2.0
    vol = nh.VolumeGeometry.cube((n, n, n), vs)
    det = nh.ConeGeometry.flat(400.0, 600.0, 160, 1.5, 140, 1.5)
    g = (torch.arange(n, dtype=torch.float32) + 0.5) * vs - n * vs / 2
    Z, Y, X = torch.meshgrid(g, g, g, indexing="ij")
    c = (30.0, -20.0, 16.0)
    mu = 0.02
    ball = (mu * (((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) <= 24.0 ** 2)).to(DEV)
    nv = 180
    ang = torch.tensor((2 * np.pi * np.arange(nv) / nv).astype(np.float32))
    sino = nh.project(ball, vol, det, ang)
    rec = F.fdk(sino, vol, det, ang).cpu().numpy()

    def region(p, half=3):
        i = [int(round(p[k] / vs + n / 2)) for k in range(3)]
        return rec[i[2] - half:i[2] + half, i[1] - half:i[1] + half, i[0] - half:i[0] + half].mean()

    assert region(c) == pytest.approx(mu, rel=0.04)
    for m in ((-c[0], c[1], c[2]), (c[0], -c[1], c[2]), (c[0], c[1], -c[2]), (c[1], c[0], c[2])):
        assert abs(region(m)) < 0.1 * mu, m
    # the backprojection stage against the oracle on the GPU's filtered views
    scale = F.fdk_scale(det, nv, 2 * math.pi / nv, False)
    q = F.ramp_filter(F.preweight(sino, det, scale=scale))
    od = oracle.det_struct(0, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    want = oracle.fdk_backproject(ov, od, ang.numpy().astype(np.float64), q.cpu().numpy().astype(np.float64))
    got = F.backproject(q, vol, det, ang).cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-5 * np.linalg.norm(want)
