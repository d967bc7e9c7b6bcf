"""Regression tests of input handling at the C ABI / operator boundary.

* Warp displacements far outside the volume (beyond 2^31 voxels) read nothing
  and produce zeros in the warp, its transpose and its field gradient.
* The fixed-point adjoint keeps full precision for sinograms of tiny
  magnitude (its scale 2^e then lies beyond the FP32 range).
* Shape mismatches between projections, detector, angles and offsets, and
  between warp inputs, raise before anything is launched.
* The voxel-driven backprojection on one-row / one-column detectors (every
  view skipped) and beyond its offset limits.
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
