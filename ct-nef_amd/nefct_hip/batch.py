"""Batched Siddon pair over many problems of uneven size.

One launch per direction covers every problem: the forward's workgroups are
16x16-ray tasks (one view, neighbouring views dealt to the same XCD) and the
adjoint's 32x32x16 voxel blocks, numbered across all problems, so small
problems do not leave the GPU idle and there is one launch instead of one per
problem (the reference loops
in Python over views and phases, nefct/functions/project_t.py:153-159,
back_project_t.py:150-156).

``SiddonBatch`` owns static device buffers (volumes, sinograms, problem table,
workspace), so a whole reconstruction iteration that uses it can be captured
in a HIP graph (``torch.cuda.CUDAGraph``) and replayed.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np
import torch

from . import _lib
from .geometry import ConeGeometry, VolumeGeometry
from .ops import _stream_ptr


class Problem(ctypes.Structure):
    """Mirror of ``nefct_problem`` (include/nefct_hip.h)."""
    _fields_ = [
        ("volume", ctypes.c_void_p), ("sinogram", ctypes.c_void_p), ("angles", ctypes.c_void_p),
        ("grid", ctypes.c_int * 3), ("center", ctypes.c_float * 3), ("size", ctypes.c_float * 3),
        ("nv", ctypes.c_int), ("cyl", ctypes.c_int), ("SID", ctypes.c_float), ("SAD", ctypes.c_float),
        ("na", ctypes.c_int), ("nb", ctypes.c_int), ("da", ctypes.c_float), ("ai", ctypes.c_float),
        ("db", ctypes.c_float), ("bi", ctypes.c_float),
        ("tile_begin", ctypes.c_longlong), ("block_begin", ctypes.c_longlong), ("ws_offset", ctypes.c_longlong),
        ("tiles_a", ctypes.c_int), ("tiles_b", ctypes.c_int), ("blocks", ctypes.c_int * 3), ("pad_", ctypes.c_int),
    ]


def _bind():
    return _lib.load()


class SiddonBatch:
    """Static batch of (volume geometry, detector, angles) problems.

    ``volumes[i]`` ([nz, ny, nx]) and ``sinograms[i]`` ([nv, nb, na]) are the
    device buffers the kernels read and write; ``project()`` fills the
    sinograms from the volumes, ``backproject()`` the volumes from the
    sinograms (exact adjoint, every voxel written)."""

    def __init__(self, problems: Sequence[tuple], device="cuda", flags: int = 0):
        L = _bind()
        self.device = torch.device(device)
        self.flags = int(flags)
        self.geoms = [(vol, det) for vol, det, _ in problems]
        self.angles = [torch.as_tensor(np.asarray(a.cpu() if torch.is_tensor(a) else a, np.float32)).to(self.device)
                       for _, _, a in problems]
        self.volumes: List[torch.Tensor] = []
        self.sinograms: List[torch.Tensor] = []
        n = len(problems)
        self.host = (Problem * n)()
        for i, (vol, det, _) in enumerate(problems):
            ang = self.angles[i]
            v = torch.zeros(vol.shape_zyx, dtype=torch.float32, device=self.device)
            s = torch.zeros((ang.numel(), det.nb, det.na), dtype=torch.float32, device=self.device)
            self.volumes.append(v)
            self.sinograms.append(s)
            p = self.host[i]
            p.volume, p.sinogram, p.angles = v.data_ptr(), s.data_ptr(), ang.data_ptr()
            p.grid = (ctypes.c_int * 3)(*vol.grid)
            p.center = (ctypes.c_float * 3)(*vol.center)
            p.size = (ctypes.c_float * 3)(*vol.size)
            p.nv, p.cyl = ang.numel(), det.cyl
            p.SID, p.SAD, p.na, p.nb = det.SID, det.SAD, det.na, det.nb
            p.da, p.ai, p.db, p.bi = det.da, det.ai, det.db, det.bi
        wsb = ctypes.c_size_t(0)
        tiles = ctypes.c_longlong(0)
        blocks = ctypes.c_longlong(0)
        _lib.check(L.nefct_batch_plan(ctypes.addressof(self.host), n, self.flags, ctypes.addressof(wsb),
                                      ctypes.addressof(tiles), ctypes.addressof(blocks)), "batch_plan")
        self.nprob = n
        self.total_tiles = tiles.value
        self.total_blocks = blocks.value
        self.ws_bytes = wsb.value
        self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        raw = bytes(ctypes.string_at(ctypes.addressof(self.host), ctypes.sizeof(self.host)))
        self.table = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
        torch.cuda.synchronize(self.device)

    @property
    def rays(self) -> int:
        return int(sum(s.numel() for s in self.sinograms))

    def project(self):
        L = _bind()
        st = L.nefct_batch_project(ctypes.addressof(self.host), ctypes.c_void_p(self.table.data_ptr()), self.nprob, self.total_tiles,
                                   self.flags, ctypes.c_void_p(self.workspace.data_ptr()), self.ws_bytes,
                                   ctypes.c_void_p(_stream_ptr(self.device)))
        _lib.check(st, "batch_project")
        return self.sinograms

    def tune(self, reps: int = 2) -> dict:
        """Times the batched adjoint with neighbouring lanes one and two
        detector rows apart (identical results) on the batch's current
        sinograms and keeps the faster for ``backproject``.  Returns the
        record (ms per call of each)."""
        base = self.flags & ~_lib.F_LANE_STRIDE2
        times = {}
        for name, fl in (("stride1", base), ("stride2", base | _lib.F_LANE_STRIDE2)):
            self.flags = fl
            self.backproject()
            torch.cuda.synchronize(self.device)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                self.backproject()
            e1.record()
            torch.cuda.synchronize(self.device)
            times[name] = round(e0.elapsed_time(e1) / reps, 3)
        best = min(times, key=times.get)
        self.flags = base | (_lib.F_LANE_STRIDE2 if best == "stride2" else 0)
        return {"adjoint": best, "ms": times}

    def backproject(self):
        L = _bind()
        st = L.nefct_batch_backproject(ctypes.addressof(self.host), ctypes.c_void_p(self.table.data_ptr()), self.nprob,
                                       self.total_blocks, self.flags, ctypes.c_void_p(self.workspace.data_ptr()),
                                       self.ws_bytes, ctypes.c_void_p(_stream_ptr(self.device)))
        _lib.check(st, "batch_backproject")
        return self.volumes


class BatchedProjection(torch.autograd.Function):
    """Differentiable batched forward: inputs are the problem volumes, outputs
    their sinograms; backward runs the batched adjoint.  Inputs are copied
    into the batch's static buffers, outputs are cloned out of them."""

    @staticmethod
    def forward(ctx, batch: SiddonBatch, *volumes):
        ctx.batch = batch
        for dst, src in zip(batch.volumes, volumes):
            dst.copy_(src)
        batch.project()
        return tuple(s.clone() for s in batch.sinograms)

    @staticmethod
    def backward(ctx, *grads):
        batch = ctx.batch
        for dst, g in zip(batch.sinograms, grads):
            if g is None:
                dst.zero_()
            else:
                dst.copy_(g)
        batch.backproject()
        return (None,) + tuple(v.clone() for v in batch.volumes)


def batched_project(batch: SiddonBatch, volumes):
    return BatchedProjection.apply(batch, *volumes)
