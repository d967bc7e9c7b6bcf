"""GPU parity of the warp and FDK kernels against the oracle (HIP path, C ABI).

Bars: warp forward/transpose/displacement gradient within 1e-6 relative L2 of
the FP64 oracle (the sample position is formed in FP32 on both sides);
FDK stages within 1e-5 relative L2; dot tests of the warp pair to 1e-6.
"""
import ctypes
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


def _field(shape, seed, amp=2.5):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-amp, amp, size=shape).astype(np.float32) for _ in range(3)]


@pytest.mark.parametrize("boundary", [0, 1])
def test_deform_vs_oracle(oracle, boundary):
    from nefct_hip.deform import deform, deform_adjoint_op, deform_dvf_grad_op
    shape = (20, 23, 70)   # ragged x (not a multiple of the 256-lane block)
    rng = np.random.default_rng(5)
    img = rng.uniform(size=shape).astype(np.float32)
    g = rng.normal(size=shape).astype(np.float32)
    mx, my, mz = _field(shape, 6)
    t = lambda a: torch.tensor(a, device=DEV)
    bname = "ref" if boundary == 0 else "zero"
    out = deform(t(img), t(mx), t(my), t(mz), boundary).cpu().numpy()
    assert _rel(out, oracle.deform(img, mx, my, mz, bname)) < 1e-6
    adj = deform_adjoint_op(t(g), t(mx), t(my), t(mz), boundary).cpu().numpy()
    assert _rel(adj, oracle.deform_adjoint(g, mx, my, mz, bname)) < 1e-6
    dg = deform_dvf_grad_op(t(img), t(g), t(mx), t(my), t(mz), boundary).cpu().numpy()
    ref = oracle.deform_dvf_grad(img, g, mx, my, mz, bname)
    for k in range(3):
        assert _rel(dg[k], ref[k]) < 1e-6


def test_deform_reference_boundary_semantics():
    """Zero displacement: interior copied, the outer shell zero (the reference
    never uses base index 0 or n-1, cpp/deform/deform.cu.cc:31-33)."""
    from nefct_hip.deform import deform
    x = torch.rand((9, 10, 11), device=DEV)
    z = torch.zeros_like(x)
    out = deform(x, z, z, z, 0)
    assert torch.equal(out[1:-1, 1:-1, 1:-1], x[1:-1, 1:-1, 1:-1])
    assert out[0].abs().max().item() == 0 and out[:, :, -1].abs().max().item() == 0
    assert torch.equal(deform(x, z, z, z, 1), x)


@pytest.mark.parametrize("boundary", [0, 1])
def test_warp_dot_test_and_autograd(boundary):
    from nefct_hip.deform import Warp, deform, deform_adjoint_op
    shape = (16, 18, 20)
    g = torch.Generator().manual_seed(3)
    x = torch.randn(shape, generator=g).to(DEV)
    y = torch.randn(shape, generator=g).to(DEV)
    dvf = (torch.rand((3,) + shape, generator=g) * 4 - 2).to(DEV)
    Wx = deform(x, dvf[0], dvf[1], dvf[2], boundary)
    WTy = deform_adjoint_op(y, dvf[0], dvf[1], dvf[2], boundary)
    lhs = (Wx.double() * y.double()).sum().item()
    rhs = (x.double() * WTy.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    xi = x.clone().requires_grad_(True)
    di = dvf.clone().requires_grad_(True)
    (Warp.apply(xi, di, boundary) * y).sum().backward()
    assert torch.allclose(xi.grad, WTy, atol=1e-6, rtol=1e-5)
    # displacement gradient by central differences inside one cell
    frac = dvf - torch.floor(dvf)                # in [0, 1), also for negative displacements
    base = torch.floor(dvf) + 0.2 + 0.6 * frac   # keep fractional parts in [0.2, 0.8]
    dv = base.clone().requires_grad_(True)
    (Warp.apply(x, dv, boundary) * y).sum().backward()
    eps = 0.1
    for k in range(3):
        e = torch.zeros_like(base)
        e[k] = eps
        fp = (deform(x, *(base + e), boundary) * y)
        fm = (deform(x, *(base - e), boundary) * y)
        fd = (fp - fm) / (2 * eps)
        assert torch.allclose(dv.grad[k], fd, atol=2e-5, rtol=2e-4), k


def _det_small():
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry.cube((40, 40, 32), 2.0)
    det = ConeGeometry.flat(300.0, 450.0, 100, 1.5, 80, 1.5)
    return vol, det


def test_fdk_stages_vs_oracle(oracle):
    from nefct_hip import fdk as F
    vol, det = _det_small()
    od = oracle.det_struct(0, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    rng = np.random.default_rng(8)
    nv = 30
    sino = rng.uniform(size=(nv, det.nb, det.na)).astype(np.float32)
    ang = (2 * np.pi * np.arange(nv) / nv).astype(np.float32)
    s = torch.tensor(sino, device=DEV)
    # cosine weight
    q = F.preweight(s, det, scale=1.0).cpu().numpy()
    assert _rel(q, sino * oracle.cosine_weight(od)[None]) < 1e-6
    # Parker on a short-scan subset
    dang = float(ang[1] - ang[0])
    qp = F.preweight(s, det, parker=True, dang_rad=dang, direction=1, scale=2.0).cpu().numpy()
    want = sino * oracle.cosine_weight(od)[None] * oracle.parker_weights(nv, dang, od)[:, None, :] * 2.0
    assert _rel(qp, want) < 1e-6
    # ramp filter rows
    fr = F.ramp_filter(s, "hann").cpu().numpy()
    assert _rel(fr, oracle.ramp_filter_rows(sino, "hann")) < 1e-5
    # backprojection
    bp = F.backproject(s, vol, det, torch.tensor(ang)).cpu().numpy()
    assert _rel(bp, oracle.fdk_backproject(ov, od, ang.astype(np.float64), sino.astype(np.float64))) < 1e-5


@pytest.mark.parametrize("na", [2, 7])
def test_fdk_backproject_narrow_detector_partial_z_chunk(oracle, na):
    """Backprojection edge cases: a detector two or seven columns wide (the
    two taps of a row read as one 8-byte load at the row's end) and nz = 13,
    which leaves a partial chunk of z voxels for the last threads."""
    from nefct_hip import ConeGeometry, VolumeGeometry
    from nefct_hip import fdk as F
    vol = VolumeGeometry.cube((9, 11, 13), 1.0)
    det = ConeGeometry.flat(60.0, 90.0, na, 3.0, 6, 4.0)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(0, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    rng = np.random.default_rng(na)
    nv = 12
    q = rng.standard_normal((nv, det.nb, det.na)).astype(np.float32)
    ang = (2 * np.pi * np.arange(nv) / nv + 0.1).astype(np.float32)
    bp = F.backproject(torch.tensor(q, device=DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    want = oracle.fdk_backproject(ov, od, ang.astype(np.float64), q.astype(np.float64))
    assert np.count_nonzero(want) > 0
    assert _rel(bp, want) < 1e-5


@pytest.mark.parametrize("na,rows", [(1, 3), (2, 1), (37, 5), (64, 4), (100, 3), (200, 5), (333, 7), (1024, 3),
                                     (2048, 2)])
def test_fdk_filter_arbitrary_taps(na, rows):
    """nefct_fdk_filter is the linear row convolution with any (2 na - 1) taps
    (centre na - 1): random non-symmetric taps, odd row counts, every FFT
    length from 64 to 4096 (radix-8 passes ending in radix 8, 4 or 2), against
    the direct sum in FP64."""
    from nefct_hip import _lib
    from nefct_hip.ops import _stream_ptr
    rng = np.random.default_rng(na * 31 + rows)
    x = rng.standard_normal((rows, na)).astype(np.float32)
    h = rng.standard_normal(2 * na - 1).astype(np.float32)
    want = np.stack([np.convolve(r.astype(np.float64), h.astype(np.float64))[na - 1:2 * na - 1] for r in x])
    xi, hi = torch.tensor(x, device=DEV), torch.tensor(h, device=DEV)
    out = torch.full_like(xi, float("nan"))
    L = _lib.load()
    vp = ctypes.c_void_p
    st = L.nefct_fdk_filter(vp(xi.data_ptr()), vp(out.data_ptr()), rows, na, vp(hi.data_ptr()),
                            vp(_stream_ptr(xi.device)))
    _lib.check(st, "fdk_filter")
    assert _rel(out.cpu().numpy(), want) < 1e-5


def test_fdk_helical_offsets_vs_oracle(oracle):
    """Per-view z offsets shift source and detector together
    (nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:44-65): equal to the
    unshifted backprojection of a volume moved by -offset."""
    from nefct_hip import fdk as F
    from nefct_hip import VolumeGeometry
    vol, det = _det_small()
    rng = np.random.default_rng(9)
    nv = 12
    sino = torch.tensor(rng.uniform(size=(nv, det.nb, det.na)).astype(np.float32), device=DEV)
    ang = torch.tensor((2 * np.pi * np.arange(nv) / nv).astype(np.float32))
    off = 4.0
    a = F.backproject(sino, vol, det, ang, offsets=torch.full((nv,), off))
    vol2 = VolumeGeometry(vol.grid, (vol.center[0], vol.center[1], vol.center[2] - off), vol.size)
    b = F.backproject(sino, vol2, det, ang)
    assert _rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-5


def test_fdk_uniform_ball_from_gpu_projections():
    """Known answer: FDK of the exact projections of a uniform ball returns its
    attenuation at the centre (full scan and Parker short scan)."""
    import nefct_hip as nh
    from nefct_hip import fdk as F
    vol = nh.VolumeGeometry.cube((64, 64, 64), 2.0)
    det = nh.ConeGeometry.flat(400.0, 600.0, 160, 1.5, 140, 1.5)
    g = (torch.arange(64, dtype=torch.float32) + 0.5) * 2.0 - 64
    Z, Y, X = torch.meshgrid(g, g, g, indexing="ij")
    mu = 0.02
    ball = (mu * ((X ** 2 + Y ** 2 + Z ** 2) <= 40.0 ** 2)).to(DEV)
    for parker in (False, True):
        if parker:
            fan = 2 * math.degrees(det.fan_half_angle())
            nv = int(math.ceil((180 + fan) / 1.0)) + 1
            ang = torch.tensor(np.radians(1.0 * np.arange(nv)).astype(np.float32))
        else:
            ang = torch.tensor((2 * np.pi * np.arange(180) / 180).astype(np.float32))
        sino = nh.project(ball, vol, det, ang)
        rec = F.fdk(sino, vol, det, ang, parker=parker)
        c = rec[28:36, 28:36, 28:36].mean().item()
        assert c == pytest.approx(mu, rel=0.03), (parker, c)
