"""Drop-in mirror of the reference's operator and caller interfaces.

Operators (TensorFlow custom ops of nefct, nefct/cc/siddon_3d/siddon.cc:11-80,
cpp/deform/deform.cc:12-22), same names, argument order and meaning:

* ``ProjectFlat / ProjectCyli(image, grid, center, size, angle, SID, SAD, na, da,
  ai, nb, db, bi)`` -> ``projection_value`` [nv, nb, na].  Reproduces the
  reference op including its 1/L^2 scaling and slab attribution (reference
  mode flags) and its passing ``da`` as ``db`` (siddon.cc:155, :211).
* ``BackProjectFlat / BackProjectCyli(projection_value, grid, center, size,
  angle, SID, SAD, na, da, ai, nb, db, bi, nx, ny, nz)`` -> ``image`` allocated
  with shape (nx, ny, nz) and filled x fastest, exactly as the reference
  allocates {nx, ny, nz} and writes image[ix + nx*(iy + ny*iz)]
  (siddon.cc:263, siddon.cu.cc:527).
* ``Deform(image, mx, my, mz, grid)`` -> warped image, reference boundary rule,
  zero-filled output (nefct/cc/deform/deform.cc:73-107).
* ``deform_tex.Deform / DeformInvert / Shift / ShiftZ / Rotate`` -- the ops of
  the texture warp library (nefct/cc/deform/deform_tex.cc:12-66): clamp-to-edge
  trilinear sampling, fixed-point inverse displacement field (30 iterations).

``grid``/``center``/``size`` may be lists or tensors (on any device); they are
read on the host once.  Inputs must be HIP-device float32 tensors whose memory
is x fastest, as the reference expects.

Callers (nefct/functions/project.py, back_project.py, project_t.py,
back_project_t.py): ``Project3D``, ``BackProject3D``, ``Project3DT``,
``Project3DNT``, ``BackProject3DT``, ``BackProject3DNT`` with the fields of the
reference classes and its data layout (numpy-order [x, y, z] images,
[a, b, view] projections).  The reference classes call an out-of-tree op
(``dist_3d_mod.so``, no source in the repository), so their arithmetic cannot
be reproduced; these mirrors call the exact Siddon pair of this package in mm
units (the reference's unit scaling by ``unit_size``, project.py:74-97, is
then the identity), centre the volume on the axis as the op does (no centre
is passed to it), and pass the per-view ``offsets`` (mm) on as the helical
lift of source and detector.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib
from . import ops
from .deform import deform_invert_tex, deform_op, deform_tex_op, rotate_tex_op, shift_tex_op
from .formats import IOMixin
from .geometry import ConeGeometry, VolumeGeometry
from .siddon import backproject, project


def _host3(v, kind):
    if torch.is_tensor(v):
        v = v.detach().cpu().tolist()
    return [kind(x) for x in list(v)]


def _flat_angles(angle, device):
    a = angle if torch.is_tensor(angle) else torch.tensor(np.asarray(angle, np.float32))
    return a.to(device=device, dtype=torch.float32).reshape(-1).contiguous()


def _project(image, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, cyl):
    g = _host3(grid, int)
    c = _host3(center, float)
    s = _host3(size, float)
    ang = _flat_angles(angle, image.device)
    # the reference op hands da to the launcher as db (siddon.cc:155, :211)
    return ops.siddon_project(image.reshape(-1), g, c, s, ang, float(SID), float(SAD), int(na), float(da),
                              float(ai), int(nb), float(da), float(bi), cyl, _lib.REF_FORWARD)


def ProjectFlat(image, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi):
    """Replaces ProjectFlatOp (siddon.cc:114-168)."""
    return _project(image, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, 0)


def ProjectCyli(image, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi):
    """Replaces ProjectCyliOp (siddon.cc:170-224)."""
    return _project(image, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, 1)


def _backproject(pv, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, nx, ny, nz, cyl):
    g = _host3(grid, int)
    c = _host3(center, float)
    s = _host3(size, float)
    ang = _flat_angles(angle, pv.device)
    vol = ops.siddon_backproject(pv.reshape(-1), g, c, s, ang, float(SID), float(SAD), int(na), float(da), float(ai),
                                 int(nb), float(db), float(bi), cyl, _lib.REF_ADJOINT)
    # the reference allocates {nx, ny, nz} and fills it x fastest (siddon.cc:263)
    return vol.reshape(int(nx), int(ny), int(nz))


def BackProjectFlat(projection_value, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, nx, ny, nz):
    """Replaces BackProjectFlatOp (siddon.cc:226-286)."""
    return _backproject(projection_value, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, nx, ny, nz, 0)


def BackProjectCyli(projection_value, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, nx, ny, nz):
    """Replaces BackProjectCyliOp (siddon.cc:288-348); flat sinogram indexing (Q5 not reproduced)."""
    return _backproject(projection_value, grid, center, size, angle, SID, SAD, na, da, ai, nb, db, bi, nx, ny, nz, 1)


def Deform(image, mx, my, mz, grid):
    """Replaces DeformOp (cpp/deform/deform.cc:59-91) with the zero-filled output
    of nefct/cc/deform/deform.cc:94."""
    g = _host3(grid, int)
    shp = image.shape
    v = lambda t: t.reshape(g[2], g[1], g[0])
    return deform_op(v(image), v(mx), v(my), v(mz), 0).reshape(shp)


class _DeformTexOps:
    """The ops of the reference's texture warp library
    (nefct/cc/deform/deform_tex.cc:12-66), same names and arguments; volumes
    are reshaped to (nz, ny, nx) from ``grid`` = (nx, ny, nz)."""

    @staticmethod
    def _v(t, g):
        return t.reshape(g[2], g[1], g[0])

    def Deform(self, image, mx, my, mz, grid):
        """Replaces the Deform op of deform_tex.cc:12-22 (deform_tex.cu.cc:29-76)."""
        g = _host3(grid, int)
        v = lambda t: self._v(t, g)
        return deform_tex_op(v(image), v(mx), v(my), v(mz)).reshape(image.shape)

    def DeformInvert(self, image, mx, my, mz, grid):
        """Replaces DeformInvert (deform_tex.cc:24-34, deform_tex.cu.cc:170-226)."""
        g = _host3(grid, int)
        v = lambda t: self._v(t, g)
        return deform_invert_tex(v(image), v(mx), v(my), v(mz)).reshape(image.shape)

    def Shift(self, image, grid, mx, my, mz):
        """Replaces Shift (deform_tex.cc:36-46, deform_tex.cu.cc:242-289)."""
        g = _host3(grid, int)
        return shift_tex_op(self._v(image, g), float(mx), float(my), float(mz)).reshape(image.shape)

    def ShiftZ(self, image, grid, mz):
        """Replaces ShiftZ (deform_tex.cc:48-56, deform_tex.cu.cc:306-353)."""
        g = _host3(grid, int)
        return shift_tex_op(self._v(image, g), 0.0, 0.0, float(mz)).reshape(image.shape)

    def Rotate(self, image, grid, cos_phi, sin_phi):
        """Replaces Rotate (deform_tex.cc:58-66, deform_tex.cu.cc:373-420)."""
        g = _host3(grid, int)
        return rotate_tex_op(self._v(image, g), float(cos_phi), float(sin_phi)).reshape(image.shape)


deform_tex = _DeformTexOps()


# --------------------------------------------------------------------------
# caller mirrors (nefct/functions/*.py)
# --------------------------------------------------------------------------
@dataclass
class DetectorConfig(IOMixin):
    """DetectorDirectAConfig / DetectorDirectBConfig (nefct/geometry/detector_config.py:8-37)."""
    number: int
    size: float
    offset: float = 0.0

    @property
    def unit_size(self):
        return self.size / self.number


@dataclass
class ScannerConfig3D(IOMixin):
    """nefct/geometry/scanner_config.py:44-77: SID = source-detector, SAD = source-axis."""
    mode: str
    SID: float
    SAD: float
    detector_a: DetectorConfig
    detector_b: DetectorConfig

    def cone(self) -> ConeGeometry:
        a, b = self.detector_a, self.detector_b
        return ConeGeometry.from_scanner_config(self.mode, self.SID, self.SAD, a.number, a.size, a.offset,
                                                b.number, b.size, b.offset)


@dataclass
class Image3D(IOMixin):
    """nefct/data/image.py:46-51: data indexed [x, y, z], box centre/size in mm."""
    data: torch.Tensor
    center: tuple
    size: tuple
    timestamps: Optional[np.ndarray] = None

    @property
    def shape(self):
        return list(self.data.shape[:3])

    @property
    def unit_size(self):
        return [s / n for s, n in zip(self.size, self.shape)]

    def geometry(self) -> VolumeGeometry:
        return VolumeGeometry(tuple(self.shape), tuple(float(c) for c in self.center),
                              tuple(float(s) for s in self.size))


@dataclass
class ProjectionSequence3D(IOMixin):
    """nefct/data/projection.py:69-113: data indexed [a, b, view]."""
    data: torch.Tensor
    scanner: ScannerConfig3D
    angles: np.ndarray
    offsets: Optional[np.ndarray] = None
    timestamps: Optional[np.ndarray] = None


def _to_zyx(data_xyz):
    return data_xyz.permute(2, 1, 0).contiguous()


def _to_xyz(data_zyx):
    return data_zyx.permute(2, 1, 0).contiguous()


def _to_abv(sino_vba):
    return sino_vba.permute(2, 1, 0).contiguous()


def _to_vba(data_abv):
    return data_abv.permute(2, 1, 0).contiguous()


def _op_volume(shape, size):
    """The volume the reference callers' op sees: they pass the shape and
    lengths but no centre (project.py:73-86, back_project.py:155-167), so the
    volume is centred on the rotation axis whatever ``Image3D.center`` says."""
    return VolumeGeometry(tuple(int(n) for n in shape), (0.0, 0.0, 0.0), tuple(float(v) for v in size))


def _offsets(offsets, idx=None):
    """Per-view z offsets (mm) for the views ``idx``, or None when absent or all zero."""
    if offsets is None:
        return None
    off = np.asarray(offsets, np.float32).reshape(-1)
    if idx is not None:
        off = off[idx]
    return off if np.any(off != 0) else None


@dataclass
class Project3D:
    """nefct/functions/project.py:63-98."""
    config: ScannerConfig3D
    angles: np.ndarray
    offsets: Optional[np.ndarray] = None

    def __call__(self, image: Image3D) -> ProjectionSequence3D:
        det = self.config.cone()
        p = project(_to_zyx(image.data), _op_volume(image.shape, image.size), det,
                    np.asarray(self.angles, np.float32), offsets=_offsets(self.offsets))
        return ProjectionSequence3D(_to_abv(p), self.config, np.asarray(self.angles), self.offsets)


@dataclass
class BackProject3D:
    """nefct/functions/back_project.py:146-179."""
    config: ScannerConfig3D
    shape: list
    unit_size: float

    def __call__(self, proj: ProjectionSequence3D) -> Image3D:
        size = [n * self.unit_size for n in self.shape]
        vol = _op_volume(self.shape, size)
        v = backproject(_to_vba(proj.data), vol, self.config.cone(), np.asarray(proj.angles, np.float32),
                        offsets=_offsets(proj.offsets))
        return Image3D(_to_xyz(v), (0.0, 0.0, 0.0), tuple(size))


@dataclass
class Project3DT:
    """nefct/functions/project_t.py:112-163: each view sees deformer(image, t_i).
    The reference warps the whole volume once per view; here views that share
    a timestamp share one warp."""
    config: ScannerConfig3D
    angles: np.ndarray
    offsets: Optional[np.ndarray]
    timestamps: np.ndarray
    deformer: Callable

    def __call__(self, image: Image3D) -> ProjectionSequence3D:
        det = self.config.cone()
        ang = np.asarray(self.angles, np.float32)
        ts = np.asarray(self.timestamps)
        vol = _op_volume(image.shape, image.size)
        out = torch.zeros((det.na, det.nb, len(ang)), dtype=torch.float32, device=image.data.device)
        for t in np.unique(ts):
            idx = np.nonzero(ts == t)[0]
            warped = self.deformer(image, t)
            data = warped.data if isinstance(warped, Image3D) else warped
            p = project(_to_zyx(data), vol, det, ang[idx], offsets=_offsets(self.offsets, idx))
            out[:, :, torch.as_tensor(idx, device=out.device)] = _to_abv(p)
        return ProjectionSequence3D(out, self.config, ang, self.offsets, ts)


@dataclass
class Project3DNT:
    """nefct/functions/project_t.py:166-215: phase i of an [x, y, z, t] image
    produces the views whose timestamp equals image.timestamps[i]."""
    config: ScannerConfig3D
    angles: np.ndarray
    offsets: Optional[np.ndarray]
    timestamps: np.ndarray

    def __call__(self, image: Image3D) -> ProjectionSequence3D:
        det = self.config.cone()
        ang = np.asarray(self.angles, np.float32)
        ts = np.asarray(self.timestamps)
        geo = _op_volume(image.data.shape[:3], image.size)
        out = torch.zeros((det.na, det.nb, len(ang)), dtype=torch.float32, device=image.data.device)
        for i, t in enumerate(np.asarray(image.timestamps)):
            idx = np.nonzero(ts == t)[0]
            if idx.size == 0:
                continue
            p = project(_to_zyx(image.data[..., i]), geo, det, ang[idx], offsets=_offsets(self.offsets, idx))
            out[:, :, torch.as_tensor(idx, device=out.device)] = _to_abv(p)
        return ProjectionSequence3D(out, self.config, ang, self.offsets, ts)


@dataclass
class BackProject3DT:
    """nefct/functions/back_project_t.py:117-159: sum over views of
    deformer(backprojection of view i, t_i, timestamps); views sharing a
    timestamp are backprojected together."""
    config: ScannerConfig3D
    shape: list
    unit_size: float
    timestamps: np.ndarray
    deformer: Callable

    def __call__(self, proj: ProjectionSequence3D) -> Image3D:
        size = [n * self.unit_size for n in self.shape]
        vol = _op_volume(self.shape, size)
        det = self.config.cone()
        ang = np.asarray(proj.angles, np.float32)
        ts = np.asarray(proj.timestamps)
        total = None
        for t in np.unique(ts):
            idx = np.nonzero(ts == t)[0]
            sub = _to_vba(proj.data[:, :, torch.as_tensor(idx, device=proj.data.device)])
            bp = _to_xyz(backproject(sub, vol, det, ang[idx], offsets=_offsets(proj.offsets, idx)))
            d = self.deformer(bp, t, self.timestamps)
            total = d if total is None else total + d
        return Image3D(total, (0.0, 0.0, 0.0), tuple(size), np.asarray(self.timestamps))


@dataclass
class BackProject3DNT:
    """nefct/functions/back_project_t.py:162-204: one backprojection per phase."""
    config: ScannerConfig3D
    shape: list
    unit_size: float
    timestamps: np.ndarray

    def __call__(self, proj: ProjectionSequence3D) -> Image3D:
        size = [n * self.unit_size for n in self.shape]
        vol = _op_volume(self.shape, size)
        det = self.config.cone()
        ang = np.asarray(proj.angles, np.float32)
        ts = np.asarray(proj.timestamps)
        phases = []
        for t in np.asarray(self.timestamps):
            idx = np.nonzero(ts == t)[0]
            if idx.size == 0:
                phases.append(torch.zeros(self.shape, dtype=torch.float32, device=proj.data.device))
                continue
            sub = _to_vba(proj.data[:, :, torch.as_tensor(idx, device=proj.data.device)])
            phases.append(_to_xyz(backproject(sub, vol, det, ang[idx], offsets=_offsets(proj.offsets, idx))))
        return Image3D(torch.stack(phases, dim=-1), (0.0, 0.0, 0.0), tuple(size), np.asarray(self.timestamps))
