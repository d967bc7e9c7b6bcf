"""GPU parity of the voxel-driven backprojector (nefct/cc/pixel_3d/pixel.cu.cc)
and of the pixel-driven footprint FDK backprojector
(nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255) against the FP64
restatements in the oracle.  Bar: 1e-5 relative L2 on smooth projections
(FP32 index arithmetic on the GPU against FP64 in the oracle)."""
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


def _smooth_sino(nv, nb, na, seed):
    rng = np.random.default_rng(seed)
    iv, ib, ia = np.meshgrid(np.arange(nv), np.arange(nb), np.arange(na), indexing="ij")
    out = np.zeros((nv, nb, na))
    for _ in range(4):
        ca, cb, sa, sb = rng.uniform(0, na), rng.uniform(0, nb), rng.uniform(3, 12), rng.uniform(3, 12)
        out += rng.uniform(0.5, 2) * np.exp(-((ia - ca) / sa) ** 2 - ((ib - cb) / sb) ** 2) * (1 + 0.1 * np.cos(iv))
    return out.astype(np.float32)


@pytest.mark.parametrize("cyl", [0, 1])
def test_pixel_backproject_vs_oracle(oracle, cyl):
    from nefct_hip.pixel import pixel_backproject_op
    na, nb, nv = 96, 70, 24
    SID, SAD = 300.0, 480.0
    da = 0.006 if cyl else 1.9
    db, ai, bi = 1.7, 0.25, -0.5
    grid, low, size = (46, 40, 30), (-35.0, -30.0, -24.0), (69.0, 60.0, 45.0)
    ang = (2 * np.pi * np.arange(nv) / nv + 0.1).astype(np.float32)
    pv = _smooth_sino(nv, nb, na, 3 + cyl)
    got = pixel_backproject_op(torch.tensor(pv, device=DEV), list(grid), list(low), list(size),
                               torch.tensor(ang, device=DEV), SID, SAD, na, da, ai, nb, db, bi, cyl)
    ref = oracle.pixel_bp(cyl, grid, low, size, ang.astype(np.float64), SID, SAD, na, da, ai, nb, db, bi,
                          pv.astype(np.float64))
    assert ref.std() > 0
    assert _rel(got.cpu().numpy(), ref) < 1e-5


def test_pixel_backproject_edges_and_ragged_shapes(oracle):
    """A detector narrower than the volume's shadow: views are skipped at the
    edges (discontinuous in the voxel position), so a few voxels may flip
    where FP32 and FP64 round a sample position across the edge; the rest
    agree to 1e-5."""
    from nefct_hip.pixel import pixel_backproject_op
    na, nb, nv = 33, 21, 9
    grid, low, size = (37, 29, 19), (-20.0, -15.0, -10.0), (40.0, 30.0, 20.0)
    ang = np.linspace(0, 5, nv).astype(np.float32)
    pv = _smooth_sino(nv, nb, na, 9)
    got = pixel_backproject_op(torch.tensor(pv, device=DEV), list(grid), list(low), list(size),
                               torch.tensor(ang, device=DEV), 200.0, 300.0, na, 1.0, 0.0, nb, 1.0, 0.0, 0)
    ref = oracle.pixel_bp(0, grid, low, size, ang.astype(np.float64), 200.0, 300.0, na, 1.0, 0.0, nb, 1.0, 0.0,
                          pv.astype(np.float64))
    g = got.cpu().numpy()
    close = np.abs(g - ref) <= 1e-5 * np.abs(ref).max()
    assert close.mean() > 0.999
    assert (ref == 0).any() and (ref > 0).any()


def test_pixel_backproject_full_size_vs_oracle(oracle):
    """512 x 512 x 256 voxels of 0.8 mm, 1024 x 768 detector of 0.6 mm, source
    1000 mm from the axis and 1500 mm from the detector, four views of uniform
    random projections, against the FP64 oracle (flat detectors take the
    LDS-staged mapping, FP64 detector columns).  Bar 2e-5: samples of
    uncorrelated neighbouring cells move with the FP32 row coordinate (~6e-5
    cell at 768 rows); measured 5.9e-6 (9.2e-6 with the op's FP32 arithmetic)."""
    from nefct_hip import ConeGeometry, VolumeGeometry
    from nefct_hip.pixel import pixel_backproject
    vol = VolumeGeometry.cube((512, 512, 256), 0.8)
    det = ConeGeometry.flat(1000.0, 1500.0, 1024, 0.6, 768, 0.6)
    ang = np.array([0.3, 0.91, 2.2, 4.0], np.float32)
    gen = torch.Generator(device=DEV).manual_seed(1)
    p = torch.rand((4, det.nb, det.na), generator=gen, device=DEV)
    got = pixel_backproject(p, vol, det, torch.tensor(ang, device=DEV)).cpu().numpy()
    low = [c - s / 2 for c, s in zip(vol.center, vol.size)]
    ref = oracle.pixel_bp(0, vol.grid, low, vol.size, ang.astype(np.float64), det.SID, det.SAD, det.na, det.da,
                          det.ai, det.nb, det.db, det.bi, p.cpu().numpy().astype(np.float64))
    assert ref.std() > 0
    assert _rel(got, ref) < 2e-5


@pytest.mark.parametrize("cyl,helical", [(0, False), (1, False), (0, True)])
def test_footprint_backproject_vs_oracle(oracle, cyl, helical):
    from nefct_hip.fdk import footprint_backproject
    nx, ny, nz = 40, 36, 24
    na, nb, nv = 64, 40, 20
    SID, SAD = 90.0, 60.0
    da = 0.02 if cyl else 1.25
    db, ai, bi = 1.1, 0.3, -0.2
    ang = (2 * np.pi * np.arange(nv) / nv + 0.05).astype(np.float32)
    off = (np.linspace(-3, 3, nv).astype(np.float32) if helical else None)
    pv = _smooth_sino(nv, nb, na, 11 + cyl)
    got = footprint_backproject(torch.tensor(pv, device=DEV), (nx, ny, nz), torch.tensor(ang, device=DEV),
                                None if off is None else torch.tensor(off, device=DEV), SID=SID, SAD=SAD, da=da,
                                ai=ai, na=na, db=db, bi=bi, nb=nb, cyl=bool(cyl))
    ref = oracle.fdk_fp_bp(cyl, (nx, ny, nz), ang.astype(np.float64), None if off is None else off.astype(np.float64),
                           SID, SAD, na, da, ai, nb, db, bi, pv.astype(np.float64))
    assert ref.std() > 0
    assert _rel(got.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("geom", ["fine", "coarse"])
def test_footprint_backproject_lds_window_paths(oracle, geom):
    """The kernel pre-sums each workgroup's footprints over chunks of slabs in
    an LDS window.  "fine": C2's proportions in voxel units (detector column
    0.5 voxel at the axis), views at and around 45 deg so workgroups whose
    pixels disagree on the walk axis add directly, most chunks fit the window;
    "coarse": 2-voxel columns, most chunks exceed it and add directly."""
    from nefct_hip.fdk import footprint_backproject
    if geom == "fine":
        nx, ny, nz, na, nb = 96, 96, 48, 192, 144
        SID, SAD, da, ai, db, bi = 375.0, 250.0, 0.75, 0.0, 0.75, 0.0
        ang = np.deg2rad([0.0, 20.0, 40.0, 45.0, 50.0, 100.0, 225.0]).astype(np.float32)
    else:
        nx, ny, nz, na, nb = 48, 40, 32, 48, 24
        SID, SAD, da, ai, db, bi = 150.0, 100.0, 3.0, 0.2, 2.5, 0.1
        ang = np.deg2rad([0.0, 30.0, 45.0, 77.0, 190.0]).astype(np.float32)
    nv = ang.size
    pv = _smooth_sino(nv, nb, na, 17)
    pv[:, :, ::7] = 0.0                      # zero pixels add nothing (and are skipped)
    got = footprint_backproject(torch.tensor(pv, device=DEV), (nx, ny, nz), torch.tensor(ang, device=DEV),
                                None, SID=SID, SAD=SAD, da=da, ai=ai, na=na, db=db, bi=bi, nb=nb)
    ref = oracle.fdk_fp_bp(0, (nx, ny, nz), ang.astype(np.float64), None, SID, SAD, na, da, ai, nb, db, bi,
                           pv.astype(np.float64))
    assert ref.std() > 0
    assert _rel(got.cpu().numpy(), ref) < 1e-5


def test_footprint_backproject_full_size_vs_oracle(oracle):
    """512 x 512 x 256 voxels, 1024 x 768 detector of 0.75-voxel pixels,
    source 1250 and detector 1875 voxels from the source-axis line: two views
    (one within 3 deg of a diagonal, where workgroups straddle the walk-axis
    switch) against the FP64 oracle.  Bar 5e-5: at these distances the FP32
    edge rays (as the reference computes them) place footprint edges to ~1e-4
    voxel, which moves boundary-cell weights by ~2e-4; the full-size error is
    ~2.4e-5."""
    from nefct_hip.fdk import footprint_backproject
    nx, ny, nz, na, nb = 512, 512, 256, 1024, 768
    kw = dict(SID=1875.0, SAD=1250.0, da=0.75, ai=0.0, na=na, db=0.75, bi=0.0, nb=nb)
    ang = np.array([0.3, 0.83], np.float32)
    gen = torch.Generator(device=DEV).manual_seed(1)
    pv = torch.rand((2, nb, na), generator=gen, device=DEV)
    got = footprint_backproject(pv, (nx, ny, nz), torch.tensor(ang, device=DEV), **kw).cpu().numpy()
    ref = oracle.fdk_fp_bp(0, (nx, ny, nz), ang.astype(np.float64), None, 1875.0, 1250.0, na, 0.75, 0.0, nb, 0.75,
                           0.0, pv.cpu().numpy().astype(np.float64))
    assert ref.std() > 0
    assert _rel(got, ref) < 5e-5


def test_footprint_backproject_accumulates_into_out():
    from nefct_hip.fdk import footprint_backproject
    pv = torch.rand((3, 8, 16), device=DEV)
    kw = dict(SID=50.0, SAD=30.0, da=1.0, ai=0.0, na=16, db=1.0, bi=0.0, nb=8)
    once = footprint_backproject(pv, (12, 12, 6), torch.tensor([0.0, 1.0, 2.0]), **kw)
    out = torch.ones((6, 12, 12), device=DEV)
    footprint_backproject(pv, (12, 12, 6), torch.tensor([0.0, 1.0, 2.0]), out=out, **kw)
    assert torch.allclose(out, once + 1, rtol=1e-6, atol=1e-6)
