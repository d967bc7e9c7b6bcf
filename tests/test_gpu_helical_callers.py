"""Helical view offsets in the Siddon pair, and the reference callers'
layout and units.

* Helical: view iv's source and detector lifted by offsets[iv] along z
  (dist_3d_back_fdk.cu.cc:46,65) is the same view of the volume lowered by
  offsets[iv].  The GPU pair with offsets is compared, view by view, with the
  FP64 oracle projecting a volume whose centre is moved down by the view's
  offset (an independent route: the oracle knows nothing of offsets), and
  the adjoint with the sum of the per-view oracle transposes; the pair with
  offsets passes the dot test.
* Callers (nefct/functions/project.py:63-98, back_project.py:146-179):
  ``Project3D`` of an [x, y, z] image gives [a, b, view] data in mm in
  which a small ball lands where the line from the source through its
  centre meets the detector described by ScannerConfig3D.positions
  (scanner_config.py:55-74), offsets included, for flat and cylindrical
  detectors; ``BackProject3D`` is its transpose in those layouts.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("cyl", [0, 1])
def test_helical_pair_vs_per_view_shifted_volume(oracle, cyl):
    import nefct_hip as nh
    from nefct_hip.phantom import ellipsoid_phantom
    vol = nh.VolumeGeometry.cube((36, 32, 28), 2.0, center=(1.0, -2.0, 3.0))
    if cyl:
        det = nh.ConeGeometry("cyli", 300.0, 480.0, 60, 0.006, 0.5, 40, 2.4, 0.5)
    else:
        det = nh.ConeGeometry.flat(300.0, 480.0, 60, 2.4, 40, 2.4)
    nv = 9
    ang = (2 * math.pi * np.arange(nv) / nv + 0.05).astype(np.float32)
    off = np.linspace(-14.0, 11.0, nv).astype(np.float32)
    img = ellipsoid_phantom(vol, 21, 5)
    od = oracle.det_struct(det.cyl, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang), offsets=off).cpu().numpy()
    y = np.random.default_rng(3).standard_normal((nv, det.nb, det.na))
    got_bp = nh.backproject(torch.tensor(y, dtype=torch.float32, device=DEV), vol, det, torch.tensor(ang),
                            offsets=torch.tensor(off)).cpu().numpy()
    want = np.zeros_like(got, dtype=np.float64)
    want_bp = np.zeros(vol.shape_zyx)
    for v in range(nv):
        ov = oracle.vol_struct(vol.grid, (vol.center[0], vol.center[1], vol.center[2] - float(off[v])), vol.size)
        a1 = np.array([ang[v]], np.float64)
        want[v] = oracle.exact_fwd(ov, od, a1, img.double().numpy())[0]
        want_bp += oracle.exact_bp(ov, od, a1, y[v:v + 1])
    assert _rel(got, want) < 1e-5
    assert _rel(got_bp, want_bp) < 1e-5
    # the offsets matter: without them the result differs
    plain = nh.project(img.to(DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    assert _rel(plain, want) > 1e-2
    # dot test of the helical pair
    g = torch.Generator().manual_seed(5)
    x = torch.randn(vol.shape_zyx, generator=g)
    lhs = float((nh.project(x.to(DEV), vol, det, torch.tensor(ang), offsets=off).double().cpu()
                 * torch.tensor(y)).sum())
    rhs = float((x.double() * torch.tensor(got_bp, dtype=torch.float64)).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_helical_reference_mode_vs_shifted_restatement(oracle):
    import nefct_hip as nh
    from nefct_hip.phantom import ellipsoid_phantom
    vol = nh.VolumeGeometry.cube((32, 32, 24), 2.0)
    det = nh.ConeGeometry.flat(300.0, 480.0, 56, 2.4, 36, 2.4)
    nv = 6
    ang = (2 * math.pi * np.arange(nv) / nv + 0.2).astype(np.float32)
    off = np.array([-9.0, -3.5, 0.0, 2.25, 6.0, 12.5], np.float32)
    img = ellipsoid_phantom(vol, 22, 5)
    od = oracle.det_struct(0, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang), nh.REF_FORWARD, offsets=off).cpu().numpy()
    for v in range(nv):
        ov = oracle.vol_struct(vol.grid, (0.0, 0.0, -float(off[v])), vol.size)
        want = oracle.ref_fwd64(ov, od, np.array([ang[v]], np.float64), img.double().numpy(),
                                flags=oracle.REF_FWD_FLAGS)[0]
        assert _rel(got[v], want) < 1e-5, v


def _meshgrid(number, size, offset):
    return (0.5 + np.arange(number)) * (size / number) - size / 2 + offset


@pytest.mark.parametrize("mode", ["flat", "cylin"])
def test_project3d_ball_lands_on_reference_detector(mode):
    """Image3D.data[x, y, z] holding a voxelised ball (radius 7 mm, mirror
    symmetric about a voxel centre) seen by a fine detector (0.3 mm cells,
    so each 2 mm voxel's shadow spans ten cells and the sampled centroid
    resolves the magnification); Project3D data [a, b, view]: at view angle 0
    the centroid of the ball's projection lies where the line from the source
    through the ball's centre meets the detector of ScannerConfig3D.positions
    (source and detector lifted together by the view's z offset, so the
    detector's own b coordinate is (z - offset) times the magnification), and
    the projection's integral over the detector, divided by the magnification
    squared, is the ball's volume (line integrals in mm).  A swapped SID/SAD
    or a sign error in an offset moves the centroid by several mm; the bar is
    0.1 mm."""
    from nefct_hip.reference import DetectorConfig, Image3D, Project3D, ScannerConfig3D
    if mode == "flat":
        da = DetectorConfig(480, 144.0, 4.5)
    else:
        da = DetectorConfig(480, 0.16, 0.01)
    db = DetectorConfig(320, 96.0, -3.0)
    cfg = ScannerConfig3D(mode, 900.0, 600.0, da, db)     # SID source-detector, SAD source-axis
    n, unit = 32, 2.0
    ix, iy, iz = 21, 9, 20
    c = (np.arange(n) + 0.5) * unit - n * unit / 2
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = c[ix], c[iy], c[iz]
    ball = ((X - x) ** 2 + (Y - y) ** 2 + (Z - z) ** 2 <= 7.0 ** 2).astype(np.float32)
    image = Image3D(torch.tensor(ball, device=DEV), (5.0, -7.0, 3.0), (n * unit,) * 3)   # the op ignores the centre
    angles = np.array([0.0, 0.4], np.float32)
    oz = 6.0
    proj = Project3D(cfg, angles, np.array([oz, 0.0], np.float32))(image)
    assert tuple(proj.data.shape) == (480, 320, 2)
    p = proj.data[:, :, 0].double().cpu().numpy()
    ma, mb = _meshgrid(da.number, da.size, da.offset), _meshgrid(db.number, db.size, db.offset)
    a_c = (p.sum(axis=1) * ma).sum() / p.sum()
    b_c = (p.sum(axis=0) * mb).sum() / p.sum()
    if mode == "flat":
        s = cfg.SID / (cfg.SAD + x)          # source (-SAD, 0, oz), detector plane x = SID - SAD
        a_hit, b_hit = y * s, (z - oz) * s
        tol_a, cell_a = 0.1, da.size / da.number
    else:
        # the detector is the cylinder of radius SID about the source's axis
        s = cfg.SID / math.hypot(cfg.SAD + x, y)
        a_hit, b_hit = math.atan2(y, cfg.SAD + x), (z - oz) * s
        tol_a, cell_a = 0.1 / cfg.SID, cfg.SID * da.size / da.number
    assert abs(a_c - a_hit) < tol_a, (a_c, a_hit)
    assert abs(b_c - b_hit) < 0.1, (b_c, b_hit)
    integral = p.sum() * cell_a * (db.size / db.number) / (s * s)
    assert integral == pytest.approx(ball.sum() * unit ** 3, rel=0.01)


def test_backproject3d_is_the_transpose_in_reference_layouts():
    from nefct_hip.reference import BackProject3D, DetectorConfig, Image3D, Project3D, ProjectionSequence3D, \
        ScannerConfig3D
    cfg = ScannerConfig3D("flat", 900.0, 600.0, DetectorConfig(70, 105.0, -2.0), DetectorConfig(50, 75.0, 1.5))
    shape, unit = [30, 26, 22], 2.0
    ang = (2 * math.pi * np.arange(7) / 7).astype(np.float32)
    off = np.linspace(-5, 5, 7).astype(np.float32)
    g = torch.Generator().manual_seed(6)
    x = torch.randn(shape, generator=g).to(DEV)
    yv = torch.randn((70, 50, 7), generator=g).to(DEV)
    px = Project3D(cfg, ang, off)(Image3D(x, (0, 0, 0), tuple(s * unit for s in shape))).data
    bp = BackProject3D(cfg, shape, unit)(ProjectionSequence3D(yv, cfg, ang, off)).data
    assert tuple(bp.shape) == tuple(shape)
    lhs = (px.double() * yv.double()).sum().item()
    rhs = (x.double() * bp.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
