"""The drop-in mirrors of the reference's op and caller interfaces
(nefct/cc/siddon_3d/siddon.cc:11-80, cpp/deform/deform.cc:12-22,
nefct/functions/project*.py, back_project*.py), on the GPU."""
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


def test_project_flat_op_reproduces_reference_call(oracle):
    """ProjectFlat(image, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi):
    reference mode, and the op passes da for db (siddon.cc:155)."""
    from nefct_hip import reference as R
    from nefct_hip.phantom import ellipsoid_phantom
    from nefct_hip import VolumeGeometry
    vol = VolumeGeometry.cube((40, 40, 32), 2.0)
    img = ellipsoid_phantom(vol, 4, 6)
    ang = (2 * math.pi * np.arange(10) / 10).astype(np.float32)
    args = dict(SID=500.0, SAD=800.0, na=70, da=2.0, ai=0.5, nb=50, db=3.0, bi=0.5)
    p = R.ProjectFlat(img.to(DEV), torch.tensor([40, 40, 32], dtype=torch.int32, device=DEV),
                      torch.tensor([0.0, 0.0, 0.0]), [80.0, 80.0, 64.0], torch.tensor(ang, device=DEV), **args)
    assert tuple(p.shape) == (10, 50, 70)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(0, 500.0, 800.0, 70, 2.0, 0.5, 50, 2.0, 0.5)   # db replaced by da
    ref = oracle.ref_fwd64(ov, od, ang.astype(np.float64), img.double().numpy(), flags=oracle.REF_FWD_FLAGS)
    assert _rel(p.cpu().numpy(), ref) < 1e-5


def test_backproject_flat_op_shape_and_memory_order(oracle):
    from nefct_hip import reference as R
    ang = (2 * math.pi * np.arange(8) / 8).astype(np.float32)
    y = np.random.default_rng(0).uniform(size=(8, 30, 40)).astype(np.float32)
    out = R.BackProjectFlat(torch.tensor(y, device=DEV), [24, 20, 16], [0, 0, 0], [48.0, 40.0, 32.0],
                            torch.tensor(ang), 400.0, 650.0, 40, 2.5, 0.5, 30, 2.5, 0.5, 24, 20, 16)
    assert tuple(out.shape) == (24, 20, 16)
    ov = oracle.vol_struct((24, 20, 16), (0, 0, 0), (48.0, 40.0, 32.0))
    od = oracle.det_struct(0, 400.0, 650.0, 40, 2.5, 0.5, 30, 2.5, 0.5)
    ref = oracle.ref_bp64(ov, od, ang.astype(np.float64), y.astype(np.float64), flags=oracle.REF_BP_FLAGS)
    # same bytes as the reference: x fastest in a {nx, ny, nz}-labelled buffer
    assert _rel(out.reshape(-1).cpu().numpy(), ref.reshape(-1)) < 1e-5


def test_deform_op(oracle):
    from nefct_hip import reference as R
    rng = np.random.default_rng(1)
    img = rng.uniform(size=(12, 14, 16)).astype(np.float32)
    m = [rng.uniform(-2, 2, size=img.shape).astype(np.float32) for _ in range(3)]
    t = lambda a: torch.tensor(a, device=DEV)
    out = R.Deform(t(img), t(m[0]), t(m[1]), t(m[2]), torch.tensor([16, 14, 12], dtype=torch.int32))
    assert _rel(out.cpu().numpy(), oracle.deform(img, *m, boundary="ref")) < 1e-6


def _scanner():
    from nefct_hip import reference as R
    return R.ScannerConfig3D("flat", 900.0, 600.0, R.DetectorConfig(64, 96.0, 0.0), R.DetectorConfig(48, 72.0, 0.0))


def test_project3d_and_backproject3d_are_adjoint():
    from nefct_hip import reference as R
    sc = _scanner()
    ang = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    g = torch.Generator().manual_seed(2)
    x = torch.randn((32, 32, 24), generator=g).to(DEV)
    img = R.Image3D(x, (0.0, 0.0, 0.0), (48.0, 48.0, 36.0))
    p = R.Project3D(sc, ang)(img)
    assert tuple(p.data.shape) == (64, 48, 12)
    y = torch.randn(p.data.shape, generator=g).to(DEV)
    b = R.BackProject3D(sc, [32, 32, 24], 1.5)(R.ProjectionSequence3D(y, sc, ang))
    lhs = (p.data.double() * y.double()).sum().item()
    rhs = (x.double() * b.data.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_project3dt_groups_views_by_timestamp():
    from nefct_hip import reference as R
    from nefct_hip.deform import deform
    sc = _scanner()
    ang = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    ts = np.array([0, 1, 2] * 4, dtype=np.float64)
    x = torch.rand((32, 32, 24), device=DEV)
    img = R.Image3D(x, (0.0, 0.0, 0.0), (48.0, 48.0, 36.0))
    calls = []

    def shift(image, t):
        calls.append(t)
        d = image.data.permute(2, 1, 0).contiguous()
        z = torch.zeros_like(d)
        return deform(d, z + float(t) * 0.5, z, z, 1).permute(2, 1, 0).contiguous()

    p = R.Project3DT(sc, ang, None, ts, shift)(img)
    assert sorted(calls) == [0.0, 1.0, 2.0]   # one warp per distinct timestamp, not per view
    for t in (0.0, 1.0, 2.0):
        idx = np.nonzero(ts == t)[0]
        ref = R.Project3D(sc, ang[idx])(R.Image3D(shift(img, t), img.center, img.size))
        got = p.data[:, :, torch.as_tensor(idx, device=DEV)]
        assert torch.allclose(got, ref.data, rtol=1e-6, atol=1e-6)


def test_project3dnt_and_backproject3dnt_adjoint():
    from nefct_hip import reference as R
    sc = _scanner()
    ang = np.linspace(0, 2 * np.pi, 15, endpoint=False)
    ts = np.array([0, 1, 2] * 5, dtype=np.float64)
    g = torch.Generator().manual_seed(5)
    x = torch.randn((32, 32, 24, 3), generator=g).to(DEV)
    img = R.Image3D(x, (0.0, 0.0, 0.0), (48.0, 48.0, 36.0), np.array([0.0, 1.0, 2.0]))
    p = R.Project3DNT(sc, ang, None, ts)(img)
    y = torch.randn(p.data.shape, generator=g).to(DEV)
    b = R.BackProject3DNT(sc, [32, 32, 24], 1.5, np.array([0.0, 1.0, 2.0]))(R.ProjectionSequence3D(y, sc, ang, None, ts))
    assert tuple(b.data.shape) == (32, 32, 24, 3)
    lhs = (p.data.double() * y.double()).sum().item()
    rhs = (x.double() * b.data.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
