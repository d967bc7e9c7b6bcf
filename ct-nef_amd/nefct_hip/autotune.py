"""Autotuner for the Siddon pair's launch shapes.

Both directions have work mappings with identical results that differ in how
work maps onto the chip (include/nefct_hip.h):

* adjoint: volume blocks of 32x32x8, 32x16x8, 64x32x8, 32x32x16, 16x16x32
  or 64x16x16 voxels, neighbouring lanes one or two detector rows apart,
  fixed-point or float LDS accumulation, or ray tiles with LDS slab windows;
* forward: ray tiles of 16x16, 32x8, 8x32 or 64x4 detector cells per
  workgroup (a wavefront is the first 64 of them) and the number of
  neighbouring views each XCD walks together (its L2 reuse).

Which is fastest depends on the volume, detector and view count, so the
tuner times the candidates on a sample of the views (the forward on a
contiguous run, since its order exploits neighbouring views; the adjoint on
an even spread) and keeps the winner per (device, direction, geometry, mode)
in a JSON cache.

Cache file: ``$NEFCT_AUTOTUNE_CACHE`` or ``~/.cache/nefct_hip/autotune.json``.
Entries are read-mostly; writes replace the file atomically.
"""
from __future__ import annotations

import json
import os
import tempfile
import threading
from typing import Dict, Optional, Tuple

import torch

from . import _lib
from .geometry import ConeGeometry, VolumeGeometry

ADJOINT_CANDIDATES: Dict[str, int] = {
    "tile32x32x16": _lib.F_TILE_32x32x16,
    "tile64x32x8": _lib.F_TILE_64x32x8,
    "tile64x16x16": _lib.F_TILE_64x16x16,
    "tile32x16x8": _lib.F_TILE_32x16x8,
    "tile32x32x8": _lib.F_TILE_32x32x8,
    "tile16x16x32": _lib.F_TILE_16x16x32,
    "tile32x32x16/stride2": _lib.F_TILE_32x32x16 | _lib.F_LANE_STRIDE2,
    "tile32x32x16/float": _lib.F_TILE_32x32x16 | _lib.F_ADJ_FLOAT_ATOMIC,
    "window": _lib.F_ADJ_WINDOW,
}

FORWARD_TILES: Dict[str, int] = {
    "16x16": _lib.F_FWD_TILE_16x16,
    "32x8": _lib.F_FWD_TILE_32x8,
    "8x32": _lib.F_FWD_TILE_8x32,
    "64x4": _lib.F_FWD_TILE_64x4,
}
# explicit view groups: the default rule depends on the view count, which the
# tuner's sample does not share with the full scan
FORWARD_VIEW_GROUPS: Dict[str, int] = {"vg4": _lib.fwd_vg(2), "vg8": _lib.fwd_vg(3), "vg16": _lib.fwd_vg(4),
                                       "vg32": _lib.fwd_vg(5)}
FORWARD_CANDIDATES: Dict[str, int] = {f"{t}/{g}": tf | gf for t, tf in FORWARD_TILES.items()
                                      for g, gf in FORWARD_VIEW_GROUPS.items()}

_lock = threading.Lock()
_cache: Optional[dict] = None


def cache_path() -> str:
    p = os.environ.get("NEFCT_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "nefct_hip", "autotune.json")


def _load() -> dict:
    global _cache
    if _cache is None:
        try:
            with open(cache_path()) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def _store(key: str, value: dict):
    with _lock:
        c = _load()
        c[key] = value
        path = cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path), prefix=".autotune")
            with os.fdopen(fd, "w") as f:
                json.dump(c, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass   # an unwritable cache only costs re-tuning


def _views_bucket(nv: int) -> int:
    b = 1
    while b < nv:
        b *= 2
    return b


def key_of(vol: VolumeGeometry, det: ConeGeometry, nv: int, mode_flags: int, device, kind: str = "adjoint") -> str:
    props = torch.cuda.get_device_properties(device)
    arch = getattr(props, "gcnArchName", props.name)
    vs = ",".join(f"{v:.4g}" for v in vol.voxel)
    return (f"{arch}|{kind}|grid={vol.grid}|vox={vs}|det={det.mode},{det.na}x{det.nb},"
            f"SID={det.SID:.4g},SAD={det.SAD:.4g},da={det.da:.4g}|views~{_views_bucket(nv)}|mode={mode_flags}")


def _time_ms(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_adjoint(vol: VolumeGeometry, det: ConeGeometry, angles, mode_flags: int = 0,
                 candidates: Optional[Dict[str, int]] = None, sample_views: int = 64, reps: int = 2,
                 force: bool = False) -> Tuple[int, dict]:
    """Pick the fastest adjoint variant for this geometry.  ``mode_flags``
    are the semantic flags (0 exact, or REF_ADJOINT) kept in every
    candidate.  Returns (flags, record); the record lists every candidate's
    time on the view sample."""
    from .siddon import backproject
    ang = torch.as_tensor(angles, dtype=torch.float32).reshape(-1)
    nv = ang.numel()
    dev = torch.device("cuda", torch.cuda.current_device())
    key = key_of(vol, det, nv, mode_flags, dev)
    if not force:
        rec = _load().get(key)
        if rec is not None:
            return int(rec["flags"]), rec
    step = max(1, nv // sample_views)
    sub = ang[::step][:sample_views].to(dev)
    g = torch.Generator(device=dev).manual_seed(0)
    y = torch.rand((sub.numel(), det.nb, det.na), generator=g, device=dev)
    cands = dict(candidates or ADJOINT_CANDIDATES)
    times = {}
    for name, fl in cands.items():
        flags = int(mode_flags) | int(fl)
        times[name] = _time_ms(lambda: backproject(y, vol, det, sub, flags=flags), reps)
    best = min(times, key=times.get)
    rec = {"flags": int(mode_flags) | int(cands[best]), "variant": best, "sample_views": int(sub.numel()),
           "ms_on_sample": {k: round(v, 3) for k, v in times.items()}}
    _store(key, rec)
    return int(rec["flags"]), rec


def tune_forward(vol: VolumeGeometry, det: ConeGeometry, angles, mode_flags: int = 0,
                 candidates: Optional[Dict[str, int]] = None, sample_views: Optional[int] = None, reps: int = 2,
                 force: bool = False) -> Tuple[int, dict]:
    """Pick the fastest forward tile and view group for this geometry, timed
    on the first ``sample_views`` views (a contiguous run, as the work order
    groups neighbouring views; default all views up to 720: on a short run the
    XCDs' shares of few view groups rank the groups differently than the full
    scan does, measured on MI355X at 512x512x256 / 1024x768: 32-view groups
    are the slowest on the first 96 views and the fastest on all 720).
    Returns (flags, record)."""
    from .siddon import project
    ang = torch.as_tensor(angles, dtype=torch.float32).reshape(-1)
    nv = ang.numel()
    dev = torch.device("cuda", torch.cuda.current_device())
    key = key_of(vol, det, nv, mode_flags, dev, "forward")
    if not force:
        rec = _load().get(key)
        if rec is not None:
            return int(rec["flags"]), rec
    sub = ang[:sample_views if sample_views else 720].to(dev)
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.rand(vol.shape_zyx, generator=g, device=dev)
    cands = dict(candidates or FORWARD_CANDIDATES)
    times = {}
    for name, fl in cands.items():
        flags = int(mode_flags) | int(fl)
        times[name] = _time_ms(lambda: project(x, vol, det, sub, flags=flags), reps)
    best = min(times, key=times.get)
    rec = {"flags": int(mode_flags) | int(cands[best]), "variant": best, "sample_views": int(sub.numel()),
           "ms_on_sample": {k: round(v, 3) for k, v in times.items()}}
    _store(key, rec)
    return int(rec["flags"]), rec


def tuned_flags(vol: VolumeGeometry, det: ConeGeometry, nv: int, mode_flags: int = 0,
                kind: str = "adjoint") -> int:
    """Flags of the cached winner for this geometry and direction ("adjoint"
    or "forward"), or the mode flags alone (default variant) when it has not
    been tuned."""
    dev = torch.device("cuda", torch.cuda.current_device())
    rec = _load().get(key_of(vol, det, nv, mode_flags, dev, kind))
    return int(rec["flags"]) if rec else int(mode_flags)


def project_tuned(image: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles, mode_flags: int = 0,
                  tune: bool = True) -> torch.Tensor:
    """``project`` with the tuned variant (tuning on first use)."""
    from .siddon import project
    if tune:
        flags, _ = tune_forward(vol, det, angles, mode_flags)
    else:
        flags = tuned_flags(vol, det, torch.as_tensor(angles).numel(), mode_flags, "forward")
    return project(image, vol, det, angles, flags=flags)


def backproject_tuned(sino: torch.Tensor, vol: VolumeGeometry, det: ConeGeometry, angles, mode_flags: int = 0,
                      tune: bool = True) -> torch.Tensor:
    """``backproject`` with the tuned variant (tuning on first use)."""
    from .siddon import backproject
    if tune:
        flags, _ = tune_adjoint(vol, det, angles, mode_flags)
    else:
        flags = tuned_flags(vol, det, torch.as_tensor(angles).numel(), mode_flags)
    return backproject(sino, vol, det, angles, flags=flags)
