"""GPU parity at the real sizes of the two plain workloads.

C1 (128^3 @ 1 mm, 256 x 256 @ 1.5 mm, 180 views, 10 ellipsoids) in full:
exact forward and adjoint against the FP64 merged-alpha oracle, reference
mode against the FP64 restatement of the reference rule, and the dot test.

C2 (512 x 512 x 256 @ 0.8 mm, 1024 x 768 @ 0.6 mm, 720 views, 30 ellipsoids):
the forward of all 720 views on the GPU, compared on every 90th view with the
oracle; the adjoint of that view subset at full volume and detector size; FDK
of the short scan compared with the oracle's backprojection on a z slab.

Bars: 1e-5 relative L2 (forward, adjoint, FDK backprojection), 1e-6 dot test.
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


def _setup(oracle, wl):
    from nefct_hip.phantom import ellipsoid_phantom
    img = ellipsoid_phantom(wl.vol, wl.seed, wl.ellipsoids)
    ov = oracle.vol_struct(wl.vol.grid, wl.vol.center, wl.vol.size)
    d = wl.det
    od = oracle.det_struct(d.cyl, d.SID, d.SAD, d.na, d.da, d.ai, d.nb, d.db, d.bi)
    return img, ov, od


@pytest.fixture(scope="module")
def c1(oracle):
    from nefct_hip import configs
    wl = configs.c1()
    img, ov, od = _setup(oracle, wl)
    return wl, img, ov, od


def test_c1_exact_forward_full(oracle, c1):
    import nefct_hip as nh
    wl, img, ov, od = c1
    ang = wl.angles()
    got = nh.project(img.to(DEV), wl.vol, wl.det, torch.tensor(ang)).cpu().numpy()
    want = oracle.exact_fwd(ov, od, ang.astype(np.float64), img.double().numpy())
    assert _rel(got, want) < 1e-5


def test_c1_exact_adjoint_full_and_dot(oracle, c1):
    import nefct_hip as nh
    wl, img, ov, od = c1
    ang = wl.angles()
    g = torch.Generator().manual_seed(11)
    y = torch.randn((wl.nv, wl.det.nb, wl.det.na), generator=g)
    got = nh.backproject(y.to(DEV), wl.vol, wl.det, torch.tensor(ang)).cpu().numpy()
    want = oracle.exact_bp(ov, od, ang.astype(np.float64), y.double().numpy())
    assert _rel(got, want) < 1e-5
    x = torch.randn(wl.vol.shape_zyx, generator=g)
    Ax = nh.project(x.to(DEV), wl.vol, wl.det, torch.tensor(ang)).double().cpu()
    lhs = float((Ax * y.double()).sum())
    rhs = float((x.double() * torch.tensor(got, dtype=torch.float64)).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_c1_reference_mode_full(oracle, c1):
    import nefct_hip as nh
    wl, img, ov, od = c1
    ang = wl.angles()
    a64 = ang.astype(np.float64)
    got = nh.project(img.to(DEV), wl.vol, wl.det, torch.tensor(ang), nh.REF_FORWARD).cpu().numpy()
    want = oracle.ref_fwd64(ov, od, a64, img.double().numpy(), flags=oracle.REF_FWD_FLAGS)
    assert _rel(got, want) < 1e-5
    y = np.random.default_rng(12).uniform(size=(wl.nv, wl.det.nb, wl.det.na)).astype(np.float32)
    got = nh.backproject(torch.tensor(y, device=DEV), wl.vol, wl.det, torch.tensor(ang), nh.REF_ADJOINT).cpu().numpy()
    want = oracle.ref_bp64(ov, od, a64, y.astype(np.float64), flags=oracle.REF_BP_FLAGS)
    assert _rel(got, want) < 1e-5


@pytest.fixture(scope="module")
def c2(oracle):
    from nefct_hip import configs
    wl = configs.c2()
    img, ov, od = _setup(oracle, wl)
    return wl, img, ov, od


C2_STRIDE = 90   # views 0, 90, ..., 630: 8 views spread over the scan


def test_c2_forward_all_views_strided_parity(oracle, c2):
    import nefct_hip as nh
    wl, img, ov, od = c2
    ang = wl.angles()
    sub = ang[::C2_STRIDE].astype(np.float64)
    x = img.to(DEV)
    got = nh.project(x, wl.vol, wl.det, torch.tensor(ang))
    assert got.shape == (720, 768, 1024)
    got = got[::C2_STRIDE].cpu().numpy()
    want = oracle.exact_fwd(ov, od, sub, img.double().numpy())
    assert _rel(got, want) < 1e-5
    got = nh.project(x, wl.vol, wl.det, torch.tensor(ang), nh.REF_FORWARD)[::C2_STRIDE].cpu().numpy()
    want = oracle.ref_fwd64(ov, od, sub, img.double().numpy(), flags=oracle.REF_FWD_FLAGS)
    assert _rel(got, want) < 1e-5


def test_c2_adjoint_view_subset_parity(oracle, c2):
    import nefct_hip as nh
    wl, img, ov, od = c2
    sub = wl.angles()[::C2_STRIDE]
    y = np.random.default_rng(13).uniform(size=(len(sub), wl.det.nb, wl.det.na)).astype(np.float32)
    yt = torch.tensor(y, device=DEV)
    got = nh.backproject(yt, wl.vol, wl.det, torch.tensor(sub)).cpu().numpy()
    want = oracle.exact_bp(ov, od, sub.astype(np.float64), y.astype(np.float64))
    assert _rel(got, want) < 1e-5
    got = nh.backproject(yt, wl.vol, wl.det, torch.tensor(sub), nh.REF_ADJOINT).cpu().numpy()
    want = oracle.ref_bp64(ov, od, sub.astype(np.float64), y.astype(np.float64), flags=oracle.REF_BP_FLAGS)
    assert _rel(got, want) < 1e-5


def test_c2_fdk_short_scan_z_slab(oracle, c2):
    """FDK of the C2 forward projection (short scan, cosine + Parker + ramp):
    the GPU backprojection of the GPU-filtered views equals the oracle's on
    four z planes through the volume; the filtered views equal the oracle's
    weighting and filter on a few views; the reconstruction is close to the
    phantom inside the fully sampled cylinder."""
    import nefct_hip as nh
    from nefct_hip import VolumeGeometry, configs
    from nefct_hip import fdk as F
    wl, img, ov, od = c2
    idx = configs.short_scan_views(wl)
    ang = wl.angles()[idx]
    sino = nh.project(img.to(DEV), wl.vol, wl.det, torch.tensor(wl.angles()))[torch.tensor(idx, device=DEV)]
    nv = len(idx)
    dang = float(ang[1] - ang[0])
    q = F.ramp_filter(F.preweight(sino, wl.det, parker=True, dang_rad=dang, direction=1,
                                  scale=F.fdk_scale(wl.det, nv, dang, True)))
    # weighting + filter on views 0, 200, nv - 1
    for v in (0, 200, nv - 1):
        s = sino[v].double().cpu().numpy()
        w = s * oracle.cosine_weight(od) * oracle.parker_weights(nv, dang, od)[v][None, :]
        w = oracle.ramp_filter_rows(w) * oracle.fdk_scale(od, nv, dang, True)
        assert _rel(q[v].cpu().numpy(), w) < 1e-5, v
    rec = F.backproject(q, wl.vol, wl.det, torch.tensor(ang))
    vz = wl.vol.voxel[2]
    qn = q.double().cpu().numpy()
    for z0 in (30, 127, 200):
        cz = wl.vol.center[2] - wl.vol.size[2] / 2 + (z0 + 2) * vz
        slab = VolumeGeometry((512, 512, 4), (wl.vol.center[0], wl.vol.center[1], cz), (wl.vol.size[0], wl.vol.size[1], 4 * vz))
        ovs = oracle.vol_struct(slab.grid, slab.center, slab.size)
        want = oracle.fdk_backproject(ovs, od, ang.astype(np.float64), qn)
        assert _rel(rec[z0:z0 + 4].cpu().numpy(), want) < 1e-5, z0
    # sanity of the whole pipeline: close to the phantom in the inner cylinder
    r = rec.cpu().numpy()
    g = (np.arange(512) + 0.5) * 0.8 - 204.8
    inner = (g[None, :] ** 2 + g[:, None] ** 2) < 150.0 ** 2
    ph = img.numpy()[100:156]
    err = np.linalg.norm((r[100:156] - ph)[:, inner]) / np.linalg.norm(ph[:, inner])
    assert err < 0.15
