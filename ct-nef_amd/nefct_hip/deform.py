"""Trilinear warp by a dense displacement field (the reference's ``Deform`` op)
and its autograd Function.

``deform(image, mx, my, mz)`` mirrors the op ``Deform(image, mx, my, mz, grid)``
(cpp/deform/deform.cc:12-22): displacements in voxel units, volumes x fastest.
``boundary=0`` reproduces the reference's rule (a voxel is written only when
every base index lies in [1, n-2], cpp/deform/deform.cu.cc:31-33, else 0 as the
zero-filled output of nefct/cc/deform/deform.cc:94); ``boundary=1`` is zero
padding.  ``Warp`` differentiates with respect to the image (transpose
scatter) and the displacement field.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .ops import _need_device, _ptr, _stream_ptr

_vp = ctypes.c_void_p


def _grid(t):
    if t.dim() != 3:
        raise _lib.NefctError(f"volumes are [nz, ny, nx]; got shape {tuple(t.shape)}")
    nz, ny, nx = t.shape
    return [int(nx), int(ny), int(nz)]


def _same_shape(ref, *others):
    for m in others:
        if m.shape != ref.shape:
            raise _lib.NefctError(f"shape {tuple(m.shape)} != {tuple(ref.shape)}")


@torch.library.custom_op("nefct::deform", mutates_args=())
def deform_op(image: torch.Tensor, mx: torch.Tensor, my: torch.Tensor, mz: torch.Tensor,
              boundary: int) -> torch.Tensor:
    _need_device(image, mx, my, mz)
    _same_shape(image, mx, my, mz)
    image, mx, my, mz = (t.contiguous() for t in (image, mx, my, mz))
    out = torch.empty_like(image)
    st = _lib.load().nefct_deform(_ptr(image), _ptr(mx), _ptr(my), _ptr(mz), _lib.int3(_grid(image)), _ptr(out),
                                  int(boundary), _vp(_stream_ptr(image.device)))
    _lib.check(st, "deform")
    return out


@deform_op.register_fake
def _(image, mx, my, mz, boundary):
    return torch.empty_like(image)


@torch.library.custom_op("nefct::deform_adjoint", mutates_args=())
def deform_adjoint_op(grad: torch.Tensor, mx: torch.Tensor, my: torch.Tensor, mz: torch.Tensor,
                      boundary: int) -> torch.Tensor:
    _need_device(grad, mx, my, mz)
    _same_shape(grad, mx, my, mz)
    grad, mx, my, mz = (t.contiguous() for t in (grad, mx, my, mz))
    out = torch.empty_like(grad)
    st = _lib.load().nefct_deform_adjoint(_ptr(grad), _ptr(mx), _ptr(my), _ptr(mz), _lib.int3(_grid(grad)),
                                          _ptr(out), int(boundary), _vp(_stream_ptr(grad.device)))
    _lib.check(st, "deform_adjoint")
    return out


@deform_adjoint_op.register_fake
def _(grad, mx, my, mz, boundary):
    return torch.empty_like(grad)


@torch.library.custom_op("nefct::deform_dvf_grad", mutates_args=())
def deform_dvf_grad_op(image: torch.Tensor, grad: torch.Tensor, mx: torch.Tensor, my: torch.Tensor,
                       mz: torch.Tensor, boundary: int) -> torch.Tensor:
    """Returns [3, *shape]: d<grad, deform(image, m)>/d(mx, my, mz)."""
    _need_device(image, grad, mx, my, mz)
    _same_shape(image, grad, mx, my, mz)
    image, grad, mx, my, mz = (t.contiguous() for t in (image, grad, mx, my, mz))
    out = torch.empty((3,) + tuple(image.shape), dtype=torch.float32, device=image.device)
    st = _lib.load().nefct_deform_dvf_grad(_ptr(image), _ptr(grad), _ptr(mx), _ptr(my), _ptr(mz),
                                           _lib.int3(_grid(image)), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                           int(boundary), _vp(_stream_ptr(image.device)))
    _lib.check(st, "deform_dvf_grad")
    return out


@deform_dvf_grad_op.register_fake
def _(image, grad, mx, my, mz, boundary):
    return image.new_empty((3,) + tuple(image.shape))


def deform(image, mx, my, mz, boundary: int = 0):
    """Warped volume (no autograd)."""
    return deform_op(image, mx, my, mz, int(boundary))


class Warp(torch.autograd.Function):
    """out = image(i + dvf(i)) with dvf [3, nz, ny, nx] = (mx, my, mz) in voxel
    units.  Backward: image gradient by the transpose scatter kernel,
    displacement gradient from the trilinear weight derivatives.  Only the
    inputs are kept for backward; the warped volume is never stored."""

    @staticmethod
    def forward(ctx, image, dvf, boundary=1):
        ctx.boundary = int(boundary)
        ctx.save_for_backward(image, dvf)
        return deform_op(image, dvf[0], dvf[1], dvf[2], int(boundary))

    @staticmethod
    def backward(ctx, g):
        image, dvf = ctx.saved_tensors
        g = g.contiguous()
        gi = gd = None
        if ctx.needs_input_grad[0]:
            gi = deform_adjoint_op(g, dvf[0], dvf[1], dvf[2], ctx.boundary)
        if ctx.needs_input_grad[1]:
            gd = deform_dvf_grad_op(image, g, dvf[0], dvf[1], dvf[2], ctx.boundary)
        return gi, gd, None


def warp(image, dvf, boundary: int = 1):
    return Warp.apply(image, dvf, boundary)


# --------------------------------------------------------------------------
# Texture-semantics warps (nefct/cc/deform/deform_tex.cc:12-66): clamp-to-edge
# trilinear sampling with the sample coordinate i + 1/2 at voxel i.
# --------------------------------------------------------------------------
@torch.library.custom_op("nefct::deform_tex", mutates_args=())
def deform_tex_op(image: torch.Tensor, mx: torch.Tensor, my: torch.Tensor, mz: torch.Tensor) -> torch.Tensor:
    _need_device(image, mx, my, mz)
    _same_shape(image, mx, my, mz)
    image, mx, my, mz = (t.contiguous() for t in (image, mx, my, mz))
    out = torch.empty_like(image)
    st = _lib.load().nefct_deform_tex(_ptr(image), _ptr(mx), _ptr(my), _ptr(mz), _lib.int3(_grid(image)),
                                      _ptr(out), _vp(_stream_ptr(image.device)))
    _lib.check(st, "deform_tex")
    return out


@deform_tex_op.register_fake
def _(image, mx, my, mz):
    return torch.empty_like(image)


@torch.library.custom_op("nefct::invert_dvf", mutates_args=())
def invert_dvf_op(mx: torch.Tensor, my: torch.Tensor, mz: torch.Tensor, iterations: int) -> torch.Tensor:
    """Returns [3, *shape]: the fixed-point inverse field."""
    _need_device(mx, my, mz)
    _same_shape(mx, my, mz)
    mx, my, mz = (t.contiguous() for t in (mx, my, mz))
    out = torch.empty((3,) + tuple(mx.shape), dtype=torch.float32, device=mx.device)
    st = _lib.load().nefct_invert_dvf(_ptr(mx), _ptr(my), _ptr(mz), _lib.int3(_grid(mx)), int(iterations),
                                      _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _vp(_stream_ptr(mx.device)))
    _lib.check(st, "invert_dvf")
    return out


@invert_dvf_op.register_fake
def _(mx, my, mz, iterations):
    return mx.new_empty((3,) + tuple(mx.shape))


@torch.library.custom_op("nefct::shift_tex", mutates_args=())
def shift_tex_op(image: torch.Tensor, mx: float, my: float, mz: float) -> torch.Tensor:
    _need_device(image)
    image = image.contiguous()
    out = torch.empty_like(image)
    st = _lib.load().nefct_shift_tex(_ptr(image), _lib.int3(_grid(image)), float(mx), float(my), float(mz),
                                     _ptr(out), _vp(_stream_ptr(image.device)))
    _lib.check(st, "shift_tex")
    return out


@shift_tex_op.register_fake
def _(image, mx, my, mz):
    return torch.empty_like(image)


@torch.library.custom_op("nefct::rotate_tex", mutates_args=())
def rotate_tex_op(image: torch.Tensor, cos_phi: float, sin_phi: float) -> torch.Tensor:
    _need_device(image)
    image = image.contiguous()
    out = torch.empty_like(image)
    st = _lib.load().nefct_rotate_tex(_ptr(image), _lib.int3(_grid(image)), float(cos_phi), float(sin_phi),
                                      _ptr(out), _vp(_stream_ptr(image.device)))
    _lib.check(st, "rotate_tex")
    return out


@rotate_tex_op.register_fake
def _(image, cos_phi, sin_phi):
    return torch.empty_like(image)


def deform_tex(image, mx, my, mz):
    """out[i] = image(i + m(i)), clamp-to-edge trilinear (deform_tex.cu.cc:15-26)."""
    return deform_tex_op(image, mx, my, mz)


def invert_dvf(mx, my, mz, iterations: int = 30):
    """Fixed-point inverse of a displacement field (deform_tex.cu.cc:78-100);
    returns [3, *shape]."""
    return invert_dvf_op(mx, my, mz, int(iterations))


def deform_invert_tex(image, mx, my, mz):
    """Warp by the inverse of m (deform_tex.cu.cc:170-226)."""
    inv = invert_dvf_op(mx, my, mz, 30)
    return deform_tex_op(image, inv[0], inv[1], inv[2])


def shift_tex(image, mx: float, my: float, mz: float):
    """out[i] = image(i - (mx, my, mz)) (deform_tex.cu.cc:228-239)."""
    return shift_tex_op(image, float(mx), float(my), float(mz))


def shift_z_tex(image, mz: float):
    """Shift along z only (deform_tex.cu.cc:291-302)."""
    return shift_tex_op(image, 0.0, 0.0, float(mz))


def rotate_tex(image, cos_phi: float, sin_phi: float):
    """In-plane rotation about (nx/2, ny/2) (deform_tex.cu.cc:355-370)."""
    return rotate_tex_op(image, float(cos_phi), float(sin_phi))
