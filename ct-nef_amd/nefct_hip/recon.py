"""Differentiable iterative reconstruction on the GPU projector pair.

The reference reconstructs with numpy loops around its projector ops:
SART (nefct/functions/sart.py:21-39), conjugate gradients on BP o P + mu TV^T TV
(conj_grad.py:28-62), anisotropic TV (total_variation.py:19-35) and its
adjoint (:125-149).  Here the same pieces run on the device:

* ``tv_diff``/``tv_diff_adjoint`` -- the reference's forward differences with a
  zero first slice, and their transpose;
* ``tv_smooth`` -- sum of sqrt(d^2 + eps^2) over those differences (a smooth
  total variation, differentiable everywhere);
* ``loss`` -- 0.5 ||A x - b||^2 + lam TV(x) through the autograd projector;
* ``gradient_descent`` -- fixed-step descent, step 1/L with L from power
  iteration on A^T A (+ the TV curvature bound);
* ``sart`` -- the reference's SART update x += lam BP((b - P x) / P 1) / BP(1).
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from .geometry import ConeGeometry, VolumeGeometry
from .siddon import backproject, forward_project, project


def tv_diff(x: torch.Tensor) -> torch.Tensor:
    """[3, *x.shape] forward differences along z, y, x with a zero first slice
    (total_variation.py:29-35, axes in this package's [z, y, x] order)."""
    out = torch.zeros((3,) + tuple(x.shape), dtype=x.dtype, device=x.device)
    out[0, 1:] = x[1:] - x[:-1]
    out[1, :, 1:] = x[:, 1:] - x[:, :-1]
    out[2, :, :, 1:] = x[:, :, 1:] - x[:, :, :-1]
    return out


def tv_diff_adjoint(d: torch.Tensor) -> torch.Tensor:
    """Transpose of tv_diff (InvertTotalVari, total_variation.py:125-149)."""
    out = torch.zeros(d.shape[1:], dtype=d.dtype, device=d.device)
    out[:-1] -= d[0, 1:]
    out[1:] += d[0, 1:]
    out[:, :-1] -= d[1, :, 1:]
    out[:, 1:] += d[1, :, 1:]
    out[:, :, :-1] -= d[2, :, :, 1:]
    out[:, :, 1:] += d[2, :, :, 1:]
    return out


def tv_smooth_torch(x: torch.Tensor, eps: float = 1e-6) -> torch.Tensor:
    """sum sqrt(d^2 + eps^2) over tv_diff(x), in torch ops (any device)."""
    d = tv_diff(x)
    return torch.sqrt(d * d + eps * eps).sum(dtype=torch.float64)


def _sq_norm_into(x: torch.Tensor, value: torch.Tensor, scale: float = 1.0):
    """value (FP64 device scalar) += scale * sum x^2, one kernel."""
    from .ops import _ptr, _stream_ptr
    x = x.contiguous()
    st = _lib.load().nefct_sq_norm(_ptr(x), x.numel(), float(scale), _ptr(value),
                                   ctypes.c_void_p(_stream_ptr(x.device)))
    _lib.check(st, "sq_norm")


def _tv_hip(x: torch.Tensor, eps: float, scale: float, want_value: bool, want_grad: bool, value=None):
    from .ops import _need_device, _ptr, _stream_ptr
    _need_device(x)
    L = _lib.load()
    x = x.contiguous()
    nz, ny, nx = x.shape
    grad = torch.empty_like(x) if want_grad else None
    if want_value and value is None:
        value = torch.zeros((), dtype=torch.float64, device=x.device)
    st = L.nefct_tv_smooth(_ptr(x), _lib.int3((nx, ny, nz)), float(eps), float(scale),
                           _ptr(grad) if grad is not None else None, _ptr(value) if value is not None else None,
                           ctypes.c_void_p(_stream_ptr(x.device)))
    _lib.check(st, "tv_smooth")
    return value, grad


class TVSmooth(torch.autograd.Function):
    """TV(x) = sum sqrt(d^2 + eps^2) with the HIP value/gradient kernel."""

    @staticmethod
    def forward(ctx, x, eps):
        ctx.save_for_backward(x)
        ctx.eps = eps
        value, _ = _tv_hip(x, eps, 1.0, True, False)
        return value

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        _, grad = _tv_hip(x, ctx.eps, 1.0, False, True)
        return (grad * g.to(grad.dtype)), None


def tv_smooth(x: torch.Tensor, eps: float = 1e-6) -> torch.Tensor:
    """Smoothed TV of a [nz, ny, nx] volume (FP64 value): the HIP kernel on a
    device tensor, torch ops elsewhere."""
    if x.is_cuda and x.dtype == torch.float32 and x.dim() == 3:
        return TVSmooth.apply(x, float(eps))
    return tv_smooth_torch(x, eps)


def tv_descent_step(x: torch.Tensor, data_grad: torch.Tensor, eps: float, lam: float, step: float) -> torch.Tensor:
    """x - step * (data_grad + lam * grad TV(x)) in one kernel."""
    from .ops import _need_device, _ptr, _stream_ptr
    _need_device(x, data_grad)
    x = x.contiguous()
    data_grad = data_grad.contiguous()
    out = torch.empty_like(x)
    nz, ny, nx = x.shape
    st = _lib.load().nefct_tv_descent_step(_ptr(x), _ptr(data_grad), _lib.int3((nx, ny, nz)), float(eps),
                                           float(lam), float(step), _ptr(out), ctypes.c_void_p(_stream_ptr(x.device)))
    _lib.check(st, "tv_descent_step")
    return out


def loss(x, b, vol: VolumeGeometry, det: ConeGeometry, angles, lam: float = 1e-3, eps: float = 1e-6):
    """0.5 ||A x - b||^2 + lam TV(x), accumulated in FP64."""
    r = forward_project(x, vol, det, angles) - b
    data = 0.5 * (r.double() * r.double()).sum()
    return data + lam * tv_smooth(x, eps) if lam else data


def power_iteration(vol, det, angles, iters: int = 10, device="cuda", seed: int = 0) -> float:
    """Largest eigenvalue of A^T A."""
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(vol.shape_zyx, generator=g).to(device)
    lam = 0.0
    for _ in range(iters):
        y = backproject(project(x, vol, det, angles), vol, det, angles)
        lam = float(torch.linalg.vector_norm(y) / torch.linalg.vector_norm(x))
        x = y / torch.linalg.vector_norm(y)
    return lam


def gradient_descent(b, vol, det, angles, steps: int = 50, lam: float = 1e-3, eps: float = 1e-6, x0=None,
                     lipschitz: float = None, callback=None, fused: bool = True, graph: bool = False):
    """x_{k+1} = x_k - grad f(x_k) / L from a zero start (or x0), f the
    ``loss``.  ``fused`` (device tensors): project, backproject, then one
    kernel for TV gradient + update; otherwise the autograd loss.  ``graph``
    (with ``fused``) captures one step in a HIP graph and replays it.
    Returns (x, loss history)."""
    dev = b.device
    x = torch.zeros(vol.shape_zyx, dtype=torch.float32, device=dev) if x0 is None else x0.clone()
    if lipschitz is None:
        lipschitz = power_iteration(vol, det, angles, device=dev)
    L = lipschitz * 1.05 + (12.0 * lam / eps if lam else 0.0)
    history = []
    if not (fused and b.is_cuda):
        for k in range(steps):
            x.requires_grad_(True)
            f = loss(x, b, vol, det, angles, lam, eps)
            (g,) = torch.autograd.grad(f, x)
            with torch.no_grad():
                x = (x - g / L).detach()
            history.append(float(f.detach()))
            if callback:
                callback(k, x, float(f.detach()))
        return x, history

    ang = torch.as_tensor(angles, dtype=torch.float32, device=dev)
    tv_eps = eps if lam else 1.0
    fval = torch.zeros((), dtype=torch.float64, device=dev)

    def step():
        # no torch reductions here: the loss value is summed by this
        # library's kernels, so the step can be captured in a graph
        r = project(x, vol, det, ang) - b
        g = backproject(r, vol, det, ang)
        fval.zero_()
        _sq_norm_into(r, fval, 0.5)
        if lam:
            _tv_hip(x, tv_eps, lam, True, False, value=fval)
        x.copy_(tv_descent_step(x, g, tv_eps, lam, 1.0 / L))

    if graph:
        s = torch.cuda.Stream(device=dev)
        s.wait_stream(torch.cuda.current_stream(dev))
        x_start = x.clone()
        with torch.cuda.stream(s):
            step()                       # warm-up outside the graph
        torch.cuda.current_stream(dev).wait_stream(s)
        x.copy_(x_start)
        g_ = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g_):
            step()
        x.copy_(x_start)
        run = g_.replay
    else:
        run = step
    for k in range(steps):
        run()
        history.append(float(fval))
        if callback:
            callback(k, x, history[-1])
    return x.clone(), history


def sart(b, vol, det, angles, iters: int = 10, relax: float = 1.0, x0=None):
    """SART as nefct/functions/sart.py:29-39: x += relax BP((b - P x) / P 1) / BP(1)."""
    dev = b.device
    x = torch.zeros(vol.shape_zyx, dtype=torch.float32, device=dev) if x0 is None else x0.clone()
    ones_v = torch.ones(vol.shape_zyx, dtype=torch.float32, device=dev)
    row = project(ones_v, vol, det, angles)
    col = backproject(torch.ones_like(row), vol, det, angles)
    for _ in range(iters):
        r = (b - project(x, vol, det, angles)) / torch.where(row > 0, row, torch.ones_like(row))
        r = torch.where(row > 0, r, torch.zeros_like(r))
        upd = backproject(r, vol, det, angles)
        x = x + relax * torch.where(col > 0, upd / col, torch.zeros_like(upd))
    return x


def conj_grad(b, vol, det, angles, n_iter: int = 10, mu: float = 0.0, d=None, x0=None):
    """Conjugate gradients on the normal equations of
    min 0.5 ||A x - b||^2 + 0.5 mu ||W x - d||^2 (W = tv_diff), i.e.
    (A^T A + mu W^T W) x = A^T b + mu W^T d, as nefct/functions/conj_grad.py:37-62
    (``d`` = None drops the TV term's data, mu = 0 the TV term).  Inner
    products in FP64."""
    dev = b.device
    x = torch.zeros(vol.shape_zyx, dtype=torch.float32, device=dev) if x0 is None else x0.clone()

    def normal(v):
        out = backproject(project(v, vol, det, angles), vol, det, angles)
        if mu:
            out = out + mu * tv_diff_adjoint(tv_diff(v))
        return out

    rhs = backproject(b, vol, det, angles)
    if mu and d is not None:
        rhs = rhs + mu * tv_diff_adjoint(d)
    r = rhs - normal(x)
    p = r.clone()
    rs_old = float((r.double() * r.double()).sum())
    for _ in range(n_iter):
        if rs_old == 0.0:
            break
        Ap = normal(p)
        alpha = rs_old / float((p.double() * Ap.double()).sum())
        x = x + alpha * p
        r = r - alpha * Ap
        rs_new = float((r.double() * r.double()).sum())
        p = r + (rs_new / rs_old) * p
        rs_old = rs_new
    return x


def finite_difference_check(x, b, vol, det, angles, voxels, eps_fd: float = 1e-3, lam: float = 1e-3,
                            eps: float = 1e-6):
    """Central differences of ``loss`` at the given flat voxel indices against
    the autograd gradient.  Returns (fd, grad) as FP64 tensors."""
    xr = x.detach().clone().requires_grad_(True)
    f = loss(xr, b, vol, det, angles, lam, eps)
    (g,) = torch.autograd.grad(f, xr)
    flat = g.reshape(-1)
    fds = []
    with torch.no_grad():
        for v in voxels:
            xp = x.detach().clone().reshape(-1)
            xm = xp.clone()
            xp[v] += eps_fd
            xm[v] -= eps_fd
            fp = loss(xp.reshape(x.shape), b, vol, det, angles, lam, eps)
            fm = loss(xm.reshape(x.shape), b, vol, det, angles, lam, eps)
            fds.append(float((fp - fm) / (2 * eps_fd)))
    return torch.tensor(fds, dtype=torch.float64), flat[torch.as_tensor(voxels)].double().cpu()
