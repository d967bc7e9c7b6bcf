"""PyTorch-ROCm custom operators over the C ABI (torch.library).

Each operator takes device tensors, allocates its output and workspace with
the caching allocator, and launches on torch's current HIP stream.  The HIP
library must be present and the tensors must live on a HIP device; there is
no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

import torch

from . import _lib

_vp = ctypes.c_void_p


def _stream_ptr(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _need_device(*tensors):
    for t in tensors:
        if not t.is_cuda:
            raise _lib.NefctError("nefct operators run on a HIP device only (got a CPU tensor)")
        if t.dtype != torch.float32:
            raise _lib.NefctError(f"expected float32, got {t.dtype}")


def _ptr(t):
    return _vp(t.data_ptr())


def _workspace(grid, nv, device):
    nbytes = _lib.load().nefct_siddon_workspace_size(_lib.int3(grid), int(nv))
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device), int(nbytes)


def _offsets(offsets, nv):
    """Per-view z offsets (device float32, nv entries) or None."""
    if offsets is None:
        return None
    _need_device(offsets)
    offsets = offsets.contiguous()
    if offsets.numel() != nv:
        raise _lib.NefctError(f"{offsets.numel()} offsets for {nv} views")
    return offsets


@torch.library.custom_op("nefct::siddon_project", mutates_args=())
def siddon_project(image: torch.Tensor, grid: List[int], center: List[float], size: List[float],
                   angles: torch.Tensor, SID: float, SAD: float, na: int, da: float, ai: float, nb: int,
                   db: float, bi: float, cyl: int, flags: int, offsets: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Siddon forward projection; image x fastest (numel nx*ny*nz); returns [nv, nb, na].
    ``offsets``: per-view z lift of source and detector (mm, helical), or None."""
    _need_device(image, angles)
    L = _lib.load()
    image = image.contiguous()
    angles = angles.contiguous()
    if image.numel() != grid[0] * grid[1] * grid[2]:
        raise _lib.NefctError(f"image has {image.numel()} elements, grid {grid}")
    nv = angles.numel()
    off = _offsets(offsets, nv)
    out = torch.empty((nv, nb, na), dtype=torch.float32, device=image.device)
    ws, nbytes = _workspace(grid, nv, image.device)
    common = (_lib.int3(grid), _lib.float3(center), _lib.float3(size), _ptr(angles))
    geo = (SID, SAD, na, da, ai, nb, db, bi)
    tail = (int(flags), _ptr(ws), nbytes, _vp(_stream_ptr(image.device)))
    if off is not None:
        st = L.nefct_project_helical(_ptr(image), *common, _ptr(off), nv, *geo, int(cyl), _ptr(out), *tail)
    else:
        fn = L.nefct_project_cyli if cyl else L.nefct_project_flat
        st = fn(_ptr(image), *common, nv, *geo, _ptr(out), *tail)
    _lib.check(st, "siddon_project")
    return out


@siddon_project.register_fake
def _(image, grid, center, size, angles, SID, SAD, na, da, ai, nb, db, bi, cyl, flags, offsets=None):
    return image.new_empty((angles.shape[0], nb, na))


@torch.library.custom_op("nefct::siddon_backproject", mutates_args=())
def siddon_backproject(sino: torch.Tensor, grid: List[int], center: List[float], size: List[float],
                       angles: torch.Tensor, SID: float, SAD: float, na: int, da: float, ai: float, nb: int,
                       db: float, bi: float, cyl: int, flags: int,
                       offsets: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Siddon adjoint; sino [nv, nb, na]; returns [nz, ny, nx] (x fastest).
    ``offsets``: per-view z lift of source and detector (mm, helical), or None."""
    _need_device(sino, angles)
    L = _lib.load()
    sino = sino.contiguous()
    angles = angles.contiguous()
    nv = angles.numel()
    if sino.numel() != nv * nb * na:
        raise _lib.NefctError(f"sinogram has {sino.numel()} elements, expected {nv}x{nb}x{na}")
    off = _offsets(offsets, nv)
    nx, ny, nz = grid
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=sino.device)
    ws, nbytes = _workspace(grid, nv, sino.device)
    common = (_lib.int3(grid), _lib.float3(center), _lib.float3(size), _ptr(angles))
    geo = (SID, SAD, na, da, ai, nb, db, bi)
    tail = (int(flags), _ptr(ws), nbytes, _vp(_stream_ptr(sino.device)))
    if off is not None:
        st = L.nefct_backproject_helical(_ptr(sino), *common, _ptr(off), nv, *geo, int(cyl), _ptr(out), *tail)
    else:
        fn = L.nefct_backproject_cyli if cyl else L.nefct_backproject_flat
        st = fn(_ptr(sino), *common, nv, *geo, _ptr(out), *tail)
    _lib.check(st, "siddon_backproject")
    return out


@siddon_backproject.register_fake
def _(sino, grid, center, size, angles, SID, SAD, na, da, ai, nb, db, bi, cyl, flags, offsets=None):
    return sino.new_empty((grid[2], grid[1], grid[0]))


@torch.library.custom_op("nefct::siddon_count", mutates_args=())
def siddon_count(angles: torch.Tensor, grid: List[int], center: List[float], size: List[float],
                 SID: float, SAD: float, na: int, da: float, ai: float, nb: int, db: float, bi: float,
                 cyl: int, flags: int) -> torch.Tensor:
    """Per-ray number of voxels with nonzero intersection length, int32 [nv, nb, na]."""
    _need_device(angles)
    L = _lib.load()
    angles = angles.contiguous()
    nv = angles.numel()
    out = torch.empty((nv, nb, na), dtype=torch.int32, device=angles.device)
    ws, nbytes = _workspace(grid, nv, angles.device)
    st = L.nefct_siddon_count(_lib.int3(grid), _lib.float3(center), _lib.float3(size), _ptr(angles), nv,
                              SID, SAD, na, da, ai, nb, db, bi, int(cyl), _ptr(out), int(flags), _ptr(ws), nbytes,
                              _vp(_stream_ptr(angles.device)))
    _lib.check(st, "siddon_count")
    return out


@siddon_count.register_fake
def _(angles, grid, center, size, SID, SAD, na, da, ai, nb, db, bi, cyl, flags):
    return angles.new_empty((angles.shape[0], nb, na), dtype=torch.int32)
