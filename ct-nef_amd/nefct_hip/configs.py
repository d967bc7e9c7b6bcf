"""The five benchmark workloads as concrete geometries.

C1 and C2 use source-to-isocentre 1000 mm and source-to-detector 1500 mm (the
values given for C1; at magnification 1.5 C2's 1024 x 0.6 mm detector covers
exactly the 409.6 mm volume width).  C3-C5 do not fix the distances: they use
SDD = 1.5 SOD and a detector pitch = voxel * 1.5 * side / cells, so the
detector covers the volume.
"""
from __future__ import annotations

from dataclasses import dataclass
import math

import numpy as np

from .geometry import ConeGeometry, VolumeGeometry


@dataclass(frozen=True)
class Workload:
    name: str
    vol: VolumeGeometry
    det: ConeGeometry
    nv: int
    ellipsoids: int
    seed: int

    def angles(self):
        return (2 * math.pi * np.arange(self.nv) / self.nv).astype(np.float32)

    @property
    def rays(self):
        return self.nv * self.det.na * self.det.nb


def c1():
    """128^3 @ 1.0 mm, 256 x 256 @ 1.5 mm, 180 views over 360 deg, 10 ellipsoids."""
    return Workload("C1 128^3 / 256x256 / 180 views", VolumeGeometry.cube((128, 128, 128), 1.0),
                    ConeGeometry.flat(1000.0, 1500.0, 256, 1.5, 256, 1.5), 180, 10, 1)


def c2():
    """512x512x256 @ 0.8 mm, 1024 x 768 @ 0.6 mm, 720 views over 360 deg, 30 ellipsoids."""
    return Workload("C2 512x512x256 / 1024x768 / 720 views", VolumeGeometry.cube((512, 512, 256), 0.8),
                    ConeGeometry.flat(1000.0, 1500.0, 1024, 0.6, 768, 0.6), 720, 30, 2)


def c3():
    """256^3 @ 1.0 mm, 512 x 384, 360 views (SOD 1000, SDD 1500, pitch covering the volume)."""
    return Workload("C3 256^3 / 512x384 / 360 views", VolumeGeometry.cube((256, 256, 256), 1.0),
                    ConeGeometry.flat(1000.0, 1500.0, 512, 0.75, 384, 1.0), 360, 10, 3)


def short_scan_views(wl: Workload):
    """Indices of the views covering 180 deg + fan angle, starting at view 0."""
    fan = 2 * wl.det.fan_half_angle()
    span = math.pi + fan
    step = 2 * math.pi / wl.nv
    n = int(math.ceil(span / step)) + 1
    return np.arange(min(n, wl.nv))


@dataclass(frozen=True)
class PhaseProblem:
    """One problem of a batched workload: geometry, its own angles, and the
    seeds of its phantom and displacement field."""
    vol: VolumeGeometry
    det: ConeGeometry
    angles: np.ndarray
    phantom_seed: int
    dvf_seed: int


def c4(seed: int = 4):
    """Ten respiratory phases of one 256x256x160 @ 1.0 mm base volume; phase
    views uniform in [40, 120], angles uniform over 360 deg; 384 x 288 detector
    (pitch 1.0 x 0.833 mm covers the volume at SDD/SOD = 1.5); each phase has
    its own Gaussian-bump displacement field (dvf_seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vol = VolumeGeometry.cube((256, 256, 160), 1.0)
    det = ConeGeometry.flat(1000.0, 1500.0, 384, 1.5 * 256 / 384, 288, 1.5 * 160 / 288)
    out = []
    for p in range(10):
        nv = int(rng.integers(40, 121))
        ang = np.sort(rng.uniform(0, 2 * math.pi, nv)).astype(np.float32)
        out.append(PhaseProblem(vol, det, ang, seed, 1000 * seed + p))
    return out


def c5(seed: int = 5, count: int = 64):
    """64 independent problems: cubic side in {64, 96, 128, 192, 256} @ 1 mm,
    views uniform in [60, 360] over 360 deg, detector (2 side)^2 at pitch 0.75 mm,
    SOD uniform in [500, 1000], SDD = 1.5 SOD; own phantom and warp field each."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for p in range(count):
        side = int(rng.choice([64, 96, 128, 192, 256]))
        nv = int(rng.integers(60, 361))
        sod = float(rng.uniform(500.0, 1000.0))
        vol = VolumeGeometry.cube((side, side, side), 1.0)
        det = ConeGeometry.flat(sod, 1.5 * sod, 2 * side, 0.75, 2 * side, 0.75)
        ang = (2 * math.pi * np.arange(nv) / nv).astype(np.float32)
        out.append(PhaseProblem(vol, det, ang, 100 * seed + p, 200 * seed + p))
    return out
