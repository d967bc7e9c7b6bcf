"""Warp-then-project over many problems (4-D CT phases, uneven batches).

Each problem p projects ``warp(base_p, dvf_p)`` with its own views.  Forward:
one warp launch per problem, then one batched projection launch for all
problems.  Backward: one batched adjoint launch, then the warp's
transpose (gradient to the base volume) and its displacement gradient per
problem.  The warped volumes live only in the batch's static buffers during
the forward; backward recomputes what it needs from the base volumes and
displacement fields, so nothing per phase is stored between the passes.

This replaces the reference's per-view loop that warps the whole volume for
every view and projects one view at a time
(nefct/functions/project_t.py:153-159, back_project_t.py:150-156).
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from .batch import SiddonBatch
from .deform import deform_adjoint_op, deform_dvf_grad_op, deform_op


class WarpProjectBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch: SiddonBatch, boundary: int, n: int, *tensors):
        bases, dvfs = tensors[:n], tensors[n:]
        ctx.batch = batch
        ctx.boundary = boundary
        ctx.n = n
        ctx.save_for_backward(*bases, *dvfs)
        for p in range(n):
            d = dvfs[p]
            batch.volumes[p].copy_(deform_op(bases[p], d[0], d[1], d[2], boundary))
        batch.project()
        return tuple(s.clone() for s in batch.sinograms)

    @staticmethod
    def backward(ctx, *grads):
        batch, n, boundary = ctx.batch, ctx.n, ctx.boundary
        saved = ctx.saved_tensors
        bases, dvfs = saved[:n], saved[n:]
        for s, g in zip(batch.sinograms, grads):
            if g is None:
                s.zero_()
            else:
                s.copy_(g)
        batch.backproject()
        gb, gd = [], []
        for p in range(n):
            gv = batch.volumes[p]
            d = dvfs[p]
            gb.append(deform_adjoint_op(gv, d[0], d[1], d[2], boundary) if ctx.needs_input_grad[3 + p] else None)
            gd.append(deform_dvf_grad_op(bases[p], gv, d[0], d[1], d[2], boundary)
                      if ctx.needs_input_grad[3 + n + p] else None)
        return (None, None, None) + tuple(gb) + tuple(gd)


def warp_project(batch: SiddonBatch, bases: Sequence[torch.Tensor], dvfs: Sequence[torch.Tensor],
                 boundary: int = 1) -> List[torch.Tensor]:
    """Sinograms of warp(bases[p], dvfs[p]) for every problem p of the batch
    (dvf [3, nz, ny, nx] in voxel units).  Differentiable in bases and dvfs."""
    n = len(bases)
    if n != batch.nprob or len(dvfs) != n:
        raise ValueError("one base volume and one displacement field per batch problem")
    return list(WarpProjectBatch.apply(batch, int(boundary), n, *bases, *dvfs))
