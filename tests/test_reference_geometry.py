"""The caller mirrors' geometry against the reference's own definition of it.

``ScannerConfig3D.positions`` (nefct/geometry/scanner_config.py:55-74) places
detector cell (ia, ib) at angle 0 at
  flat:        (SID - SAD, meshgrid_a[ia], meshgrid_b[ib])
  cylindrical: (SID cos(meshgrid_a[ia]) - SAD, SID sin(meshgrid_a[ia]), meshgrid_b[ib])
with meshgrid = (0.5 + k) unit - size / 2 + offset (detector_config.py:18-21)
and the source SAD from the axis; SID there is source-to-detector.  The
mirror's ``ScannerConfig3D.cone()`` converts this to the siddon-op geometry
of the kernels; the oracle's ``ray_points`` (the end points the FP64 oracle
and the kernels use) must then land on those positions, and rotating by the
view angle must rotate them about z.  A swapped SID/SAD, a sign error in the
offsets or a half-cell shift fails here.  CPU only.
"""
import math

import numpy as np
import pytest


def _meshgrid(number, size, offset):
    return (0.5 + np.arange(number)) * (size / number) - size / 2 + offset


@pytest.mark.parametrize("mode", ["flat", "cylin"])
def test_mirror_geometry_lands_on_reference_positions(oracle, mode):
    from nefct_hip.reference import DetectorConfig, ScannerConfig3D
    if mode == "flat":
        da = DetectorConfig(64, 96.0, 7.5)        # mm
    else:
        da = DetectorConfig(64, 0.8, 0.05)        # radians
    db = DetectorConfig(48, 72.0, -4.0)
    cfg = ScannerConfig3D(mode, 1500.0, 1000.0, da, db)   # SID = source-detector, SAD = source-axis
    g = cfg.cone()
    od = oracle.det_struct(g.cyl, g.SID, g.SAD, g.na, g.da, g.ai, g.nb, g.db, g.bi)
    ma = _meshgrid(da.number, da.size, da.offset)
    mb = _meshgrid(db.number, db.size, db.offset)
    for ia, ib in [(0, 0), (63, 47), (17, 30), (40, 5)]:
        if mode == "flat":
            want = np.array([cfg.SID - cfg.SAD, ma[ia], mb[ib]])
        else:
            want = np.array([cfg.SID * math.cos(ma[ia]) - cfg.SAD, cfg.SID * math.sin(ma[ia]), mb[ib]])
        S, D = oracle.ray_points(od, 0.0, ia, ib)
        assert np.allclose(S, [-cfg.SAD, 0.0, 0.0], atol=1e-9)
        assert np.allclose(D, want, atol=1e-9), (ia, ib)
        # a view at angle t sees the same cell rotated by t about z
        t = 0.7
        c, s = math.cos(t), math.sin(t)
        rot = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
        S2, D2 = oracle.ray_points(od, t, ia, ib)
        assert np.allclose(S2, rot @ np.array([-cfg.SAD, 0.0, 0.0]), atol=1e-9)
        assert np.allclose(D2, rot @ want, atol=1e-9)


def test_offsets_shift_cells_by_whole_cells(oracle):
    """An offset of k cells moves every cell centre by k cells: cell i with
    offset k da sits where cell i + k sits without it."""
    from nefct_hip.reference import DetectorConfig, ScannerConfig3D
    a0 = ScannerConfig3D("flat", 1200.0, 800.0, DetectorConfig(40, 60.0), DetectorConfig(30, 45.0)).cone()
    a3 = ScannerConfig3D("flat", 1200.0, 800.0, DetectorConfig(40, 60.0, 3 * 1.5),
                         DetectorConfig(30, 45.0, -2 * 1.5)).cone()
    d0 = oracle.det_struct(0, a0.SID, a0.SAD, a0.na, a0.da, a0.ai, a0.nb, a0.db, a0.bi)
    d3 = oracle.det_struct(0, a3.SID, a3.SAD, a3.na, a3.da, a3.ai, a3.nb, a3.db, a3.bi)
    for ia, ib in [(0, 5), (10, 20), (30, 27)]:
        _, D3 = oracle.ray_points(d3, 0.3, ia, ib)
        _, D0 = oracle.ray_points(d0, 0.3, ia + 3, ib - 2)
        assert np.allclose(D3, D0, atol=1e-9)
