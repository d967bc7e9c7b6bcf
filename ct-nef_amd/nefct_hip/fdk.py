"""FDK reconstruction (flat detector) on the GPU: pre-weighting, ramp filtering,
voxel-driven weighted backprojection.

Follows the reference's MATLAB FDK (mat_codes/filtering.m, ParkerWeight.m,
backprojection.m) and the distance weight of
nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:154, in the siddon-op geometry
(SID = source-isocentre, SAD = source-detector):

1. q = p * SDD / sqrt(SDD^2 + a^2 + b^2) [* Parker(view, column)] * scale
   (filtering.m:10-18), scale = 1/(2 da) * dbeta/2 * SDD/SOD with
   dbeta = 2 pi / nv for a full scan or 2 * dang (radians) for a Parker short
   scan (filtering.m:43-47);
2. each detector row is convolved with the spatial form of the windowed ramp
   (filtering.m:20-41,54-88): zero padding to n = max(64, 2^ceil(log2(2 na)))
   makes the FFT filter a linear convolution with taps h[(u - v) mod n]; the
   kernel takes those 2 na - 1 taps and applies them with its own FFT (length
   the smallest power of two >= 2 na - 1, two rows per complex transform,
   radix-8 Stockham passes in LDS);
3. every voxel sums (SOD/U)^2 times the bilinear sample of the filtered view at
   its projection (backprojection.m:11-25).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .geometry import ConeGeometry, VolumeGeometry
from .ops import _need_device, _ptr, _stream_ptr

_vp = ctypes.c_void_p
_TAPS = {}


def ramp_response(na: int, window: str = "ram-lak", d: float = 1.0) -> np.ndarray:
    """Frequency response of length n = max(64, 2^nextpow2(2 na)) (filtering.m:20-24,54-88)."""
    n = max(64, int(2 ** math.ceil(math.log2(2 * na))))
    k = np.arange(-(n // 2), n // 2)
    h = np.zeros(n)
    h[n // 2] = 0.25
    odd = (k % 2) == 1
    h[odd] = -1.0 / (np.pi * k[odd]) ** 2
    resp = np.abs(np.fft.fft(h)) * 2.0
    half = resp[: n // 2 + 1].copy()
    w = 2 * np.pi * np.arange(half.size) / n
    win = window.lower()
    if win == "shepp-logan":
        half[1:] *= np.sin(w[1:] / (2 * d)) / (w[1:] / (2 * d))
    elif win == "cosine":
        half[1:] *= np.cos(w[1:] / (2 * d))
    elif win == "hamming":
        half[1:] *= 0.54 + 0.46 * np.cos(w[1:] / d)
    elif win == "hann":
        half[1:] *= (1 + np.cos(w[1:] / d)) / 2
    elif win != "ram-lak":
        raise ValueError(f"unknown window {window!r}")
    half[w > np.pi * d] = 0.0
    return np.concatenate([half, half[-2:0:-1]])


def ramp_taps(na: int, window: str = "ram-lak") -> np.ndarray:
    """taps[k] = h[(k - (na - 1)) mod n], k = 0 .. 2 na - 2, h = real(ifft(response))."""
    key = (na, window.lower())
    if key not in _TAPS:
        resp = ramp_response(na, window)
        n = resp.size
        h = np.real(np.fft.ifft(resp))
        k = np.arange(2 * na - 1) - (na - 1)
        _TAPS[key] = h[k % n].astype(np.float32)
    return _TAPS[key]


def fdk_scale(det: ConeGeometry, nv: int, dang_rad: float, parker: bool) -> float:
    dbeta = 2.0 * dang_rad if parker else 2.0 * math.pi / nv
    return 1.0 / (2.0 * det.da) * dbeta / 2.0 * (det.SAD / det.SID)


def preweight(sino: torch.Tensor, det: ConeGeometry, parker: bool = False, dang_rad: float = 0.0,
              direction: int = 1, scale: float = 1.0) -> torch.Tensor:
    _need_device(sino)
    if det.cyl:
        raise _lib.NefctError("FDK weighting is implemented for flat detectors only")
    sino = sino.contiguous()
    if sino.dim() != 3 or tuple(sino.shape[1:]) != (det.nb, det.na):
        raise _lib.NefctError(f"projections have shape {tuple(sino.shape)}, expected (nv, {det.nb}, {det.na})")
    nv = sino.shape[0]
    out = torch.empty_like(sino)
    st = _lib.load().nefct_fdk_preweight(_ptr(sino), _ptr(out), nv, det.SID, det.SAD, det.na, det.da, det.ai,
                                         det.nb, det.db, det.bi, int(bool(parker)), math.degrees(dang_rad),
                                         int(direction), float(scale), _vp(_stream_ptr(sino.device)))
    _lib.check(st, "fdk_preweight")
    return out


def ramp_filter(q: torch.Tensor, window: str = "ram-lak") -> torch.Tensor:
    """Filter every detector row (last axis) with the windowed ramp (no scaling)."""
    _need_device(q)
    q = q.contiguous()
    na = q.shape[-1]
    taps = torch.from_numpy(ramp_taps(na, window)).to(q.device)
    out = torch.empty_like(q)
    rows = q.numel() // na
    st = _lib.load().nefct_fdk_filter(_ptr(q), _ptr(out), rows, na, _ptr(taps), _vp(_stream_ptr(q.device)))
    _lib.check(st, "fdk_filter")
    return out


MAPPINGS = {"auto": 0, "gather": 1, "staged": 2, "wide": 3}


def backproject(filtered: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles,
                offsets=None, mapping: str = "auto", out: torch.Tensor = None) -> torch.Tensor:
    """Voxel-driven weighted backprojection.  ``mapping`` picks the work
    mapping ("auto", "gather", "staged" (16x16x16 blocks) or "wide"
    (64x64x16 blocks); identical images).  ``out``: an existing
    [nz, ny, nx] float32 tensor on the projections' device to write every
    voxel of (else a new one)."""
    if mapping not in MAPPINGS:
        raise _lib.NefctError(f"unknown mapping {mapping!r}; one of {sorted(MAPPINGS)}")
    _need_device(filtered)
    filtered = filtered.contiguous()
    dev = filtered.device
    ang = angles if torch.is_tensor(angles) else torch.tensor(np.asarray(angles, np.float32))
    ang = ang.to(device=dev, dtype=torch.float32).contiguous()
    nv = ang.numel()
    if filtered.numel() != nv * det.nb * det.na:
        raise _lib.NefctError(f"projections have {filtered.numel()} elements, expected {nv}x{det.nb}x{det.na}")
    off = None
    if offsets is not None:
        off = (offsets if torch.is_tensor(offsets) else torch.tensor(np.asarray(offsets, np.float32)))
        off = off.to(device=dev, dtype=torch.float32).contiguous()
        if off.numel() != nv:
            raise _lib.NefctError(f"{off.numel()} offsets for {nv} views")
    L = _lib.load()
    wsb = int(L.nefct_fdk_workspace_size(nv))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    nx, ny, nz = vol.grid
    if out is None:
        img = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    else:
        if (tuple(out.shape) != (nz, ny, nx) or out.dtype != torch.float32 or out.device != dev
                or not out.is_contiguous()):
            raise _lib.NefctError(f"out must be a contiguous float32 [{nz}, {ny}, {nx}] tensor on {dev}")
        img = out
    st = L.nefct_fdk_backproject_mapped(_ptr(filtered), _lib.int3(vol.grid), _lib.float3(vol.center),
                                        _lib.float3(vol.size), _ptr(ang), _ptr(off) if off is not None else None,
                                        nv, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi,
                                        _ptr(img), MAPPINGS[mapping], _ptr(ws), wsb, _vp(_stream_ptr(dev)))
    _lib.check(st, "fdk_backproject")
    return img


def fdk(sino: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles, parker: bool = False,
        window: str = "ram-lak", offsets=None, direction: int = 1) -> torch.Tensor:
    """FDK reconstruction of ``sino`` [nv, nb, na].  With ``parker`` the views
    must be a short scan of evenly spaced angles (the Parker weights place view
    i at ((180/dang - nv)/2 + i + 1) dang degrees, ParkerWeight.m:9)."""
    ang = np.asarray(angles.cpu() if torch.is_tensor(angles) else angles, np.float64)
    nv = len(ang)
    if sino.dim() != 3 or sino.shape[0] != nv:
        raise _lib.NefctError(f"{nv} angles for projections of shape {tuple(sino.shape)}")
    dang = abs(float(ang[1] - ang[0])) if nv > 1 else 2 * math.pi
    if parker and nv > 1:
        direction = 1 if ang[1] > ang[0] else -1
    scale = fdk_scale(det, nv, dang, parker)
    q = preweight(sino, det, parker=parker, dang_rad=dang, direction=direction, scale=scale)
    q = ramp_filter(q, window)
    return backproject(q, vol, det, angles, offsets)


def footprint_backproject(filtered: torch.Tensor, shape, angles, offsets=None, *, SID: float, SAD: float,
                          da: float, ai: float, na: int, db: float, bi: float, nb: int, cyl: bool = False,
                          out: torch.Tensor = None) -> torch.Tensor:
    """Pixel-driven FDK backprojection with slab footprints, the arithmetic of
    nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255 with that file's units
    and naming: lengths in voxel units, volume centred on the axis,
    SAD = source-to-axis and SID = source-to-detector distance.
    ``filtered`` [nv, nb, na]; ``shape`` = (nx, ny, nz); ``offsets`` [nv] or
    None.  Adds into ``out`` ([nz, ny, nx], zero-filled when not given), as
    the reference launcher does."""
    _need_device(filtered)
    L = _lib.load()
    filtered = filtered.contiguous()
    ang = torch.as_tensor(angles, dtype=torch.float32, device=filtered.device).contiguous()
    nv = ang.numel()
    if filtered.numel() != nv * nb * na:
        raise _lib.NefctError(f"projections have {filtered.numel()} elements, expected {nv}x{nb}x{na}")
    nx, ny, nz = (int(v) for v in shape)
    if out is None:
        out = torch.zeros((nz, ny, nx), dtype=torch.float32, device=filtered.device)
    elif out.shape != (nz, ny, nx) or not out.is_contiguous() or out.dtype != torch.float32:
        raise _lib.NefctError("out must be a contiguous float32 [nz, ny, nx] tensor")
    off = None
    if offsets is not None:
        off = torch.as_tensor(offsets, dtype=torch.float32, device=filtered.device).contiguous()
        if off.numel() != nv:
            raise _lib.NefctError("one offset per view")
    fn = L.nefct_fdk_footprint_backproject_cyli if cyl else L.nefct_fdk_footprint_backproject_flat
    st = fn(_ptr(filtered), _lib.int3((nx, ny, nz)), None if off is None else _ptr(off), _ptr(ang), nv,
            float(SID), float(SAD), float(da), float(ai), int(na), float(db), float(bi), int(nb), _ptr(out),
            ctypes.c_void_p(_stream_ptr(filtered.device)))
    _lib.check(st, "fdk_footprint_backproject")
    return out
