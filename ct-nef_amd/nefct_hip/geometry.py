"""One statement of the cone-beam geometry, and converters from the reference's forms.

The kernels use the reference Siddon op's conventions
(nefct/cc/siddon_3d/siddon.cu.cc:341-353, :376-387):

* source ``S = (-SID cos t, -SID sin t, 0)``: SID is source-to-isocentre;
* flat detector cell ``(ia, ib)`` at local ``(SAD - SID, a, b)``, cylindrical at
  ``(SAD cos a - SID, SAD sin a, b)``, rotated by ``t`` about z: SAD is
  source-to-detector;
* ``a = (ai + ia) da - na da / 2`` and ``b = (bi + ib) db - nb db / 2``: ``ai``/``bi``
  are fractional-cell offsets (0.5 puts rays through cell centres);
* the volume is centred on ``center`` with extent ``size`` (mm), ``grid`` voxels,
  stored x fastest (torch shape ``[nz, ny, nx]``);
* sinograms are ``[nv, nb, na]``.

Other files of the reference name the two distances the other way round:
the FDK kernel (nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:44-54, SAD =
source-isocentre, SID = source-detector, physical ``ai``) and the Python
``ScannerConfig3D`` (nefct/geometry/scanner_config.py:44-77, SID = source-detector
1500, SAD = source-axis 1000, nefct/geometry/existing_geometry.py:11-12).
``ConeGeometry.from_scanner_config`` converts the latter.
"""
from __future__ import annotations

from dataclasses import dataclass
import math


@dataclass(frozen=True)
class VolumeGeometry:
    grid: tuple      # (nx, ny, nz)
    center: tuple    # mm
    size: tuple      # mm

    @property
    def voxel(self):
        return tuple(s / n for s, n in zip(self.size, self.grid))

    @property
    def shape_zyx(self):
        nx, ny, nz = self.grid
        return (nz, ny, nx)

    @staticmethod
    def cube(grid, voxel_mm, center=(0.0, 0.0, 0.0)):
        g = tuple(int(v) for v in grid)
        vs = voxel_mm if isinstance(voxel_mm, (tuple, list)) else (voxel_mm,) * 3
        return VolumeGeometry(g, tuple(float(c) for c in center), tuple(n * v for n, v in zip(g, vs)))


@dataclass(frozen=True)
class ConeGeometry:
    mode: str        # 'flat' or 'cyli'
    SID: float       # source -> isocentre (mm)
    SAD: float       # source -> detector (mm)
    na: int
    da: float        # mm (flat) or rad (cyli)
    ai: float        # fractional cell offset
    nb: int
    db: float        # mm
    bi: float

    @property
    def sod(self):
        return self.SID

    @property
    def sdd(self):
        return self.SAD

    @property
    def cyl(self):
        return 0 if self.mode.startswith("f") else 1

    @staticmethod
    def flat(sod, sdd, na, da, nb, db, ai=0.5, bi=0.5):
        return ConeGeometry("flat", float(sod), float(sdd), int(na), float(da), float(ai), int(nb), float(db),
                            float(bi))

    @staticmethod
    def from_scanner_config(mode, SID_src_det, SAD_src_axis, na, size_a, offset_a, nb, size_b, offset_b):
        """From ScannerConfig3D + DetectorDirect{A,B}Config fields
        (scanner_config.py:44-77, detector_config.py:8-37): swaps the distance
        names and turns the physical detector offsets into cell fractions
        (cell centres at ``(0.5 + k) * unit - size/2 + offset``)."""
        da = size_a / na
        db = size_b / nb
        ai = 0.5 + offset_a / da
        bi = 0.5 + offset_b / db
        return ConeGeometry("flat" if mode.startswith("f") else "cyli", float(SAD_src_axis), float(SID_src_det),
                            int(na), float(da), float(ai), int(nb), float(db), float(bi))

    def fan_half_angle(self):
        if self.cyl:
            return abs(self.na * self.da / 2)
        return math.atan(self.na * self.da / 2 / self.SAD)
