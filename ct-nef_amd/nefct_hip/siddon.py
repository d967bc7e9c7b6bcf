"""Cone-beam Siddon projector pair: plain calls and the autograd Function.

``project``/``backproject`` default to the exact Siddon pair (flags 0), which
are transposes of each other.  ``flags=REF_FORWARD`` / ``REF_ADJOINT``
reproduce the reference op's arithmetic (nefct/cc/siddon_3d/siddon.cu.cc):
1/L^2 forward scaling (Q1), its slab attribution (Q2), the short-ray skip
(Q3) and tie rule (Q6) of its adjoint.
"""
from __future__ import annotations

import torch

from . import _lib, ops
from ._lib import REF_ADJOINT, REF_FORWARD  # noqa: F401  (re-exported)
from .geometry import ConeGeometry, VolumeGeometry


def _args(vol: VolumeGeometry, det: ConeGeometry):
    return ([int(g) for g in vol.grid], [float(c) for c in vol.center], [float(s) for s in vol.size])


def _angles(angles, device):
    if not torch.is_tensor(angles):
        angles = torch.tensor(angles, dtype=torch.float32)
    return angles.to(device=device, dtype=torch.float32).contiguous()


def _offsets(offsets, device):
    if offsets is None:
        return None
    return _angles(offsets, device)


def project(image: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles, flags: int = 0,
            offsets=None) -> torch.Tensor:
    """Line integrals [nv, nb, na] of an x-fastest volume [nz, ny, nx].
    ``offsets`` (mm, one per view): helical lift of source and detector along z."""
    g, c, s = _args(vol, det)
    a = _angles(angles, image.device)
    return ops.siddon_project(image, g, c, s, a, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db,
                              det.bi, det.cyl, int(flags), _offsets(offsets, image.device))


def backproject(sino: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles, flags: int = 0,
                offsets=None) -> torch.Tensor:
    """Adjoint of ``project`` (exact transpose for flags 0): [nv, nb, na] -> [nz, ny, nx]."""
    g, c, s = _args(vol, det)
    a = _angles(angles, sino.device)
    return ops.siddon_backproject(sino, g, c, s, a, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db,
                                  det.bi, det.cyl, int(flags), _offsets(offsets, sino.device))


def count_intersections(vol: VolumeGeometry, det: ConeGeometry, angles, device="cuda") -> torch.Tensor:
    """Per-ray number of voxels with nonzero exact length (int32 [nv, nb, na])."""
    g, c, s = _args(vol, det)
    a = _angles(angles, device)
    return ops.siddon_count(a, g, c, s, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi,
                            det.cyl, 0)


# launch-shape flags: they change how the work maps onto the chip, not the
# exact pair's result (up to summation order for the float-atomic adjoint)
SHAPE_FLAGS = (_lib.F_TILE_MASK | _lib.F_LANE_STRIDE2 | _lib.F_FWD_TILE_MASK | _lib.F_FWD_VG_MASK
               | _lib.F_ADJ_FLOAT_ATOMIC | _lib.F_ADJ_WINDOW | _lib.F_DETERMINISTIC)


def _shape_flags(flags: int) -> int:
    flags = int(flags)
    if flags & ~SHAPE_FLAGS:
        raise _lib.NefctError(f"autograd projections take launch-shape flags only, got {flags:#x}")
    return flags


class SiddonProjection(torch.autograd.Function):
    """y = A x with A the exact Siddon projector; backward applies A^T with the
    hand-written adjoint kernel.  ``flags`` holds launch-shape flags only."""

    @staticmethod
    def forward(ctx, image, vol, det, angles, offsets=None, flags=0):
        ctx.geom = (vol, det, offsets, flags)
        ctx.save_for_backward(angles)
        return project(image, vol, det, angles, flags, offsets)

    @staticmethod
    def backward(ctx, grad):
        (angles,) = ctx.saved_tensors
        vol, det, offsets, flags = ctx.geom
        gi = backproject(grad.contiguous(), vol, det, angles, flags, offsets) if ctx.needs_input_grad[0] else None
        return gi, None, None, None, None, None


class SiddonBackprojection(torch.autograd.Function):
    """x = A^T y; backward applies A.  ``flags`` holds launch-shape flags only."""

    @staticmethod
    def forward(ctx, sino, vol, det, angles, offsets=None, flags=0):
        ctx.geom = (vol, det, offsets, flags)
        ctx.save_for_backward(angles)
        return backproject(sino, vol, det, angles, flags, offsets)

    @staticmethod
    def backward(ctx, grad):
        (angles,) = ctx.saved_tensors
        vol, det, offsets, flags = ctx.geom
        gs = project(grad.contiguous(), vol, det, angles, flags, offsets) if ctx.needs_input_grad[0] else None
        return gs, None, None, None, None, None


def forward_project(image, vol, det, angles, offsets=None, flags: int = 0):
    """Differentiable exact projection; ``flags`` may select launch shapes
    (e.g. an autotuned choice), never the reference rule."""
    a = _angles(angles, image.device)
    return SiddonProjection.apply(image, vol, det, a, _offsets(offsets, image.device), _shape_flags(flags))


def back_project(sino, vol, det, angles, offsets=None, flags: int = 0):
    a = _angles(angles, sino.device)
    return SiddonBackprojection.apply(sino, vol, det, a, _offsets(offsets, sino.device), _shape_flags(flags))
