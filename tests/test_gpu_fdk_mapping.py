"""The FDK backprojection's work mappings on the GPU (C ABI
nefct_fdk_backproject_mapped): per-voxel gathers and LDS-staged block
footprints (16x16x16 and 64x64x16 blocks) run the same per-voxel arithmetic in
the same view order, so their images are bitwise identical.  Checked at the C2 geometry (512^3-class volume,
1024x768 detector) on a view subset, with helical offsets, on ragged shapes
with a shifted volume and odd detector widths, on a fine detector whose block
footprints overflow the LDS buffer (the per-view fallback to gathers), and
against the FP64 oracle with each mapping.  A volume reaching behind the source
is refused by the staged mapping and handled by the automatic one."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _run(q, vol, det, ang, offsets, mapping):
    """One backprojection into a NaN-filled image: a voxel the mapping does not
    write stays NaN (a fresh tensor may hold an earlier call's image)."""
    from nefct_hip import fdk as F
    nx, ny, nz = vol.grid
    out = torch.full((nz, ny, nx), float("nan"), device=DEV)
    return F.backproject(q, vol, det, ang, offsets=offsets, mapping=mapping, out=out)


def _both(q, vol, det, ang, offsets=None):
    """gather, staged (small blocks) and automatic images; the wide-block
    image is checked against the gathered one here as well."""
    g = _run(q, vol, det, ang, offsets, "gather")
    s = _run(q, vol, det, ang, offsets, "staged")
    w = _run(q, vol, det, ang, offsets, "wide")
    a = _run(q, vol, det, ang, offsets, "auto")
    torch.cuda.synchronize()
    assert not torch.isnan(g).any(), "gathered image not fully written"
    assert torch.equal(g, w), "wide-block mapping differs from the gathered image"
    return g, s, a


def test_c2_geometry_view_subset_bitwise():
    from nefct_hip.configs import c2
    wl = c2()
    ang_all = wl.angles()
    idx = np.arange(0, len(ang_all), 12)                 # 60 views over the full turn
    ang = torch.tensor(ang_all[idx].astype(np.float32))
    gen = torch.Generator(device=DEV).manual_seed(3)
    q = torch.randn((len(idx), wl.det.nb, wl.det.na), generator=gen, device=DEV)
    g, s, a = _both(q, wl.vol, wl.det, ang)
    assert torch.count_nonzero(g).item() > 0.9 * g.numel()
    assert torch.equal(g, s) and torch.equal(g, a)


def test_helical_offsets_bitwise():
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry((96, 80, 72), (3.0, -2.0, 5.0), (96 * 1.1, 80 * 1.1, 72 * 0.9))
    det = ConeGeometry.flat(500.0, 800.0, 220, 1.0, 150, 1.2, ai=0.25, bi=-0.4)
    nv = 41
    rng = np.random.default_rng(5)
    q = torch.tensor(rng.standard_normal((nv, det.nb, det.na)).astype(np.float32), device=DEV)
    ang = torch.tensor((2 * np.pi * np.arange(nv) / nv + 0.3).astype(np.float32))
    off = torch.tensor(rng.uniform(-30.0, 30.0, nv).astype(np.float32), device=DEV)
    g, s, a = _both(q, vol, det, ang, offsets=off)
    assert torch.count_nonzero(g).item() > 0
    assert torch.equal(g, s) and torch.equal(g, a)


@pytest.mark.parametrize("grid,na,nb", [((37, 29, 21), 45, 33), ((16, 17, 15), 3, 2), ((5, 40, 33), 61, 64)])
def test_ragged_shapes_bitwise(grid, na, nb):
    """Blocks cut by the volume's edges (x, y and the z chunk), odd detector
    widths (the footprint's row length made even by widening it) and a
    two-row detector."""
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry(grid, (4.0, -6.0, 2.5), (grid[0] * 1.7, grid[1] * 1.3, grid[2] * 1.9))
    det = ConeGeometry.flat(150.0, 260.0, na, 160.0 / na, nb, 120.0 / nb, ai=0.1, bi=0.3)
    nv = 17
    rng = np.random.default_rng(grid[0] + na)
    q = torch.tensor(rng.standard_normal((nv, det.nb, det.na)).astype(np.float32), device=DEV)
    ang = torch.tensor((2 * np.pi * np.arange(nv) / nv + 0.05).astype(np.float32))
    g, s, a = _both(q, vol, det, ang)
    assert torch.count_nonzero(g).item() > 0
    assert torch.equal(g, s) and torch.equal(g, a)


def test_footprint_overflow_falls_back_bitwise():
    """A detector so fine that a block's footprint exceeds the LDS buffer at
    every view (16 voxels of 4 mm magnified 1.5x span ~120 columns of 0.25 mm
    by ~400 rows): those views are gathered, with the same result."""
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry.cube((40, 40, 32), 4.0)
    det = ConeGeometry.flat(400.0, 600.0, 1024, 0.25, 900, 0.25)
    nv = 9
    gen = torch.Generator(device=DEV).manual_seed(4)
    q = torch.randn((nv, det.nb, det.na), generator=gen, device=DEV)
    ang = torch.tensor((2 * np.pi * np.arange(nv) / nv).astype(np.float32))
    g, s, a = _both(q, vol, det, ang)
    assert torch.count_nonzero(g).item() > 0
    assert torch.equal(g, s) and torch.equal(g, a)


@pytest.mark.parametrize("mapping", ["gather", "staged", "wide"])
def test_each_mapping_vs_oracle(oracle, mapping):
    from nefct_hip import ConeGeometry, VolumeGeometry
    from nefct_hip import fdk as F
    vol = VolumeGeometry((40, 36, 30), (2.0, -1.0, 3.0), (80.0, 72.0, 60.0))
    det = ConeGeometry.flat(300.0, 450.0, 101, 1.5, 80, 1.5, ai=0.2, bi=-0.1)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(0, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    nv = 24
    rng = np.random.default_rng(11)
    q = rng.standard_normal((nv, det.nb, det.na)).astype(np.float32)
    ang = (2 * np.pi * np.arange(nv) / nv + 0.2).astype(np.float32)
    bp = _run(torch.tensor(q, device=DEV), vol, det, torch.tensor(ang), None, mapping).cpu().numpy()
    want = oracle.fdk_backproject(ov, od, ang.astype(np.float64), q.astype(np.float64))
    assert np.count_nonzero(want) > 0
    assert _rel(bp, want) < 1e-5


def test_volume_behind_source_refused_by_staged():
    from nefct_hip import _lib, ConeGeometry, VolumeGeometry
    from nefct_hip import fdk as F
    vol = VolumeGeometry.cube((32, 32, 8), 20.0)        # corners 450 mm from the axis
    det = ConeGeometry.flat(400.0, 600.0, 64, 4.0, 16, 4.0)
    nv = 6
    q = torch.ones((nv, det.nb, det.na), device=DEV)
    ang = torch.tensor((2 * np.pi * np.arange(nv) / nv).astype(np.float32))
    for m in ("staged", "wide"):
        with pytest.raises(_lib.NefctError):
            F.backproject(q, vol, det, ang, mapping=m)
    g = F.backproject(q, vol, det, ang, mapping="gather")
    a = F.backproject(q, vol, det, ang)
    assert torch.equal(g, a)
