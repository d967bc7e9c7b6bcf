"""Regression tests of input handling at the C ABI / operator boundary.

* Warp displacements far outside the volume (beyond 2^31 voxels) read nothing
  and produce zeros in the warp, its transpose and its field gradient.
* The fixed-point adjoint keeps full precision for sinograms of tiny
  magnitude (its scale 2^e then lies beyond the FP32 range).
* Shape mismatches between projections, detector, angles and offsets, and
  between warp inputs, raise before anything is launched.
* The voxel-driven backprojection on one-row / one-column detectors (every
  view skipped) and beyond its offset limits.
* The footprint backprojection's direct-add fallbacks (non-finite values,
  voxels at or behind the source's plane).
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("boundary", [0, 1])
@pytest.mark.parametrize("m", [3.0e9, -3.0e9, 2.5e10])
def test_deform_huge_displacement_is_zero(boundary, m):
    from nefct_hip.deform import deform, deform_adjoint_op, deform_dvf_grad_op
    shape = (6, 7, 9)
    g = torch.Generator().manual_seed(1)
    x = torch.rand(shape, generator=g).to(DEV)
    y = torch.randn(shape, generator=g).to(DEV)
    z = torch.zeros(shape, device=DEV)
    big = torch.full(shape, m, device=DEV)
    for f in ((big, z, z), (z, big, z), (z, z, big)):
        assert deform(x, *f, boundary).abs().max().item() == 0.0
        assert deform_adjoint_op(y, *f, boundary).abs().max().item() == 0.0
        assert deform_dvf_grad_op(x, y, *f, boundary).abs().max().item() == 0.0
    # the same call with a small field is nonzero (the zeros above are not vacuous)
    assert deform(x, z, z, z, boundary).abs().max().item() > 0.0


def test_fixed_point_adjoint_tiny_sinogram():
    import nefct_hip as nh
    from nefct_hip._lib import F_ADJ_FLOAT_ATOMIC
    vol = nh.VolumeGeometry.cube((24, 24, 16), 2.0)
    det = nh.ConeGeometry.flat(400.0, 600.0, 40, 2.0, 28, 2.0)
    ang = torch.tensor((2 * math.pi * np.arange(8) / 8).astype(np.float32))
    g = torch.Generator().manual_seed(7)
    y = torch.randn((8, det.nb, det.na), generator=g).to(DEV)
    ref = nh.backproject(y, vol, det, ang, 0).double()
    for scale in (1e-32, 1e-36):
        fix = nh.backproject(y * scale, vol, det, ang, 0).double() / scale
        flt = nh.backproject(y * scale, vol, det, ang, F_ADJ_FLOAT_ATOMIC).double() / scale
        assert torch.linalg.norm(fix - ref).item() <= 1e-6 * torch.linalg.norm(ref).item(), scale
        assert torch.linalg.norm(flt - ref).item() <= 1e-5 * torch.linalg.norm(ref).item(), scale


def test_fdk_shape_checks():
    import nefct_hip as nh
    from nefct_hip import fdk as F
    from nefct_hip._lib import NefctError
    vol = nh.VolumeGeometry.cube((16, 16, 8), 2.0)
    det = nh.ConeGeometry.flat(300.0, 450.0, 32, 1.5, 12, 1.5)
    ang = (2 * math.pi * np.arange(6) / 6).astype(np.float32)
    good = torch.zeros((6, det.nb, det.na), device=DEV)
    F.fdk(good, vol, det, ang)   # the well-formed call runs
    with pytest.raises(NefctError):
        F.preweight(torch.zeros((6, det.nb + 1, det.na), device=DEV), det)
    with pytest.raises(NefctError):
        F.backproject(torch.zeros((5, det.nb, det.na), device=DEV), vol, det, ang)
    with pytest.raises(NefctError):
        F.backproject(good, vol, det, ang, offsets=np.zeros(5, np.float32))
    with pytest.raises(NefctError):
        F.fdk(torch.zeros((5, det.nb, det.na), device=DEV), vol, det, ang)


def test_deform_shape_checks():
    from nefct_hip._lib import NefctError
    from nefct_hip.deform import deform_adjoint_op, deform_dvf_grad_op, deform_op
    a = torch.zeros((4, 5, 6), device=DEV)
    b = torch.zeros((4, 5, 5), device=DEV)
    with pytest.raises(NefctError):
        deform_adjoint_op(a, a, b, a, 1)
    with pytest.raises(NefctError):
        deform_dvf_grad_op(a, b, a, a, a, 1)
    with pytest.raises(NefctError):
        deform_op(a[None], a[None], a[None], a[None], 1)


def test_pixel_backproject_degenerate_detectors_and_limits():
    """One-row and one-column detectors skip every view (the 2 x 2 tap
    neighbourhood never fits), so the image is zero and no tap is read past the
    projection; a detector beyond the 24-bit row offsets is refused."""
    from nefct_hip import _lib
    from nefct_hip.pixel import pixel_backproject_op
    grid, low, size = [12, 10, 8], [-6.0, -5.0, -4.0], [12.0, 10.0, 8.0]
    ang = torch.tensor([0.0, 0.7, 2.0], device=DEV)
    for na, nb in ((16, 1), (1, 16)):
        pv = torch.rand((3, nb, na), device=DEV) + 1.0
        out = pixel_backproject_op(pv, grid, low, size, ang, 100.0, 150.0, na, 1.0, 0.0, nb, 1.0, 0.0, 0)
        torch.cuda.synchronize()
        assert out.abs().max().item() == 0.0, (na, nb)
    pv = torch.rand((3, 4, 4), device=DEV)
    img = torch.zeros((8, 10, 12), device=DEV)
    st = _lib.load().nefct_pixel_backproject_flat(pv.data_ptr(), _lib.int3(grid), _lib.float3(low),
                                                  _lib.float3(size), ang.data_ptr(), 3, 100.0, 150.0, 1 << 22, 1.0,
                                                  0.0, 4, 1.0, 0.0, img.data_ptr(), None)
    assert st != 0
    st = _lib.load().nefct_pixel_backproject_flat(pv.data_ptr(), _lib.int3(grid), _lib.float3(low),
                                                  _lib.float3(size), ang.data_ptr(), 3, 100.0, 150.0, 4, 1.0,
                                                  0.0, 4, 1.0, 0.0, img.data_ptr(), None)
    torch.cuda.synchronize()
    assert st == 0


def test_footprint_backproject_direct_fallbacks(oracle):
    """The footprint backprojection sums a workgroup's additions in fixed point
    only when the scale is defined: a workgroup holding a non-finite value adds
    directly (the NaN reaches the voxels its pixel touches, as with the
    reference's float atomics, and nothing else changes), and so does every
    workgroup when voxels reach the source's plane (checked against the FP64
    oracle)."""
    from nefct_hip.fdk import footprint_backproject
    nx, ny, nz, na, nb = 48, 40, 24, 96, 40
    kw = dict(SID=300.0, SAD=200.0, da=0.75, ai=0.0, na=na, db=0.75, bi=0.0, nb=nb)
    ang = torch.tensor([0.2, 1.3], device=DEV)
    gen = torch.Generator(device=DEV).manual_seed(3)
    pv = torch.rand((2, nb, na), generator=gen, device=DEV) + 0.5
    clean = pv.clone()
    clean[0, 20, 47] = 0.0
    bad = clean.clone()
    bad[0, 20, 47] = float("nan")
    ref = footprint_backproject(clean, (nx, ny, nz), ang, **kw)
    got = footprint_backproject(bad, (nx, ny, nz), ang, **kw)
    torch.cuda.synchronize()
    nan = torch.isnan(got)
    assert 0 < int(nan.sum()) < 0.01 * got.numel()
    assert torch.allclose(got[~nan], ref[~nan], rtol=1e-5, atol=1e-5 * ref.abs().max().item())
    # source 20 voxels from the axis of a 48 x 40 volume, views along the axes:
    # voxel centres on both sides of the source's plane, none closer than 0.5
    kw2 = dict(SID=60.0, SAD=20.0, da=0.75, ai=0.0, na=na, db=0.75, bi=0.0, nb=nb)
    ang2 = np.array([0.0, np.pi / 2], np.float32)
    p2 = _smooth_projections(2, nb, na)
    got2 = footprint_backproject(torch.tensor(p2, device=DEV), (nx, ny, nz), torch.tensor(ang2, device=DEV),
                                 **kw2).cpu().numpy()
    want = oracle.fdk_fp_bp(0, (nx, ny, nz), ang2.astype(np.float64), None, 60.0, 20.0, na, 0.75, 0.0, nb, 0.75,
                            0.0, p2.astype(np.float64))
    assert np.abs(want).max() > 0
    assert np.linalg.norm(got2 - want) <= 1e-5 * np.linalg.norm(want)


def _smooth_projections(nv, nb, na):
    ib, ia = np.meshgrid(np.arange(nb), np.arange(na), indexing="ij")
    one = np.exp(-((ia - na / 2) / (na / 5)) ** 2 - ((ib - nb / 2) / (nb / 4)) ** 2)
    return np.stack([one * (1 + 0.2 * v) for v in range(nv)]).astype(np.float32)
