"""Voxel-driven backprojection without weights (the reference's pixel_3d op
pair, nefct/cc/pixel_3d/pixel.cc:11-47 and pixel.cu.cc:12-204) as a
PyTorch-ROCm custom operator over the C ABI.

The op keeps the reference's own conventions, which differ from the Siddon
pair's: voxel centres at ``(i + 0.5) size / n + center`` (``center`` is the
low corner here, not the middle) and detector index
``y / d + n // 2 - offset - 0.5``.  ``VolumeGeometry`` describes volumes by
their middle, so ``pixel_backproject`` passes ``center - size / 2`` to keep
the voxel centres where the Siddon pair puts them; ``pixel_backproject_op``
takes the reference's raw arguments.
"""
from __future__ import annotations

from typing import List

import torch

from . import _lib
from .geometry import ConeGeometry, VolumeGeometry
from .ops import _need_device, _ptr, _stream_ptr, _vp


@torch.library.custom_op("nefct::pixel_backproject", mutates_args=())
def pixel_backproject_op(sino: torch.Tensor, grid: List[int], center: List[float], size: List[float],
                         angles: torch.Tensor, SID: float, SAD: float, na: int, da: float, ai: float, nb: int,
                         db: float, bi: float, cyl: int) -> torch.Tensor:
    """sino [nv, nb, na] -> image [nz, ny, nx]; arguments as the reference op's."""
    _need_device(sino, angles)
    L = _lib.load()
    sino = sino.contiguous()
    angles = angles.contiguous()
    nv = angles.numel()
    if sino.numel() != nv * nb * na:
        raise _lib.NefctError(f"sinogram has {sino.numel()} elements, expected {nv}x{nb}x{na}")
    nx, ny, nz = grid
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=sino.device)
    fn = L.nefct_pixel_backproject_cyli if cyl else L.nefct_pixel_backproject_flat
    st = fn(_ptr(sino), _lib.int3(grid), _lib.float3(center), _lib.float3(size), _ptr(angles), nv, SID, SAD, na,
            da, ai, nb, db, bi, _ptr(out), _vp(_stream_ptr(sino.device)))
    _lib.check(st, "pixel_backproject")
    return out


@pixel_backproject_op.register_fake
def _(sino, grid, center, size, angles, SID, SAD, na, da, ai, nb, db, bi, cyl):
    return sino.new_empty((grid[2], grid[1], grid[0]))


def pixel_backproject(sino: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles) -> torch.Tensor:
    """Voxel-driven, unweighted backprojection of ``sino`` onto ``vol``."""
    ang = torch.as_tensor(angles, dtype=torch.float32, device=sino.device)
    low = [c - s / 2 for c, s in zip(vol.center, vol.size)]
    return pixel_backproject_op(sino, list(vol.grid), low, list(vol.size), ang, det.SID, det.SAD, det.na, det.da,
                                det.ai, det.nb, det.db, det.bi, det.cyl)
