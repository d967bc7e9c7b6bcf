"""Differentiable reconstruction on the GPU pair: TV transpose, central finite
differences of the loss, monotone descent, SART (nefct/functions/sart.py)."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _setup(n=32, nv=40):
    import nefct_hip as nh
    from nefct_hip.phantom import ellipsoid_phantom
    vol = nh.VolumeGeometry.cube((n, n, n), 2.0)
    det = nh.ConeGeometry.flat(300.0, 450.0, 2 * n, 1.5 * 2.0 * n / (2 * n), 2 * n, 1.5)
    ang = torch.tensor((2 * math.pi * np.arange(nv) / nv).astype(np.float32))
    truth = ellipsoid_phantom(vol, 3, 6).to(DEV)
    b = nh.project(truth, vol, det, ang)
    return vol, det, ang, truth, b


def test_tv_adjoint_dot_gpu():
    from nefct_hip.recon import tv_diff, tv_diff_adjoint
    g = torch.Generator().manual_seed(1)
    x = torch.randn((9, 10, 11), generator=g, dtype=torch.float64).to(DEV)
    d = torch.randn((3, 9, 10, 11), generator=g, dtype=torch.float64).to(DEV)
    assert (tv_diff(x) * d).sum().item() == pytest.approx((x * tv_diff_adjoint(d)).sum().item(), rel=1e-12)


def test_loss_gradient_central_differences():
    from nefct_hip.recon import finite_difference_check
    vol, det, ang, truth, b = _setup()
    g = torch.Generator().manual_seed(2)
    x = (truth + 0.002 * torch.rand(truth.shape, generator=g).to(DEV)).contiguous()
    vox = torch.randint(0, x.numel(), (16,), generator=g).tolist()
    fd, grad = finite_difference_check(x, b, vol, det, ang, vox, eps_fd=1e-3, lam=1e-3, eps=1e-3)
    assert torch.allclose(fd, grad, rtol=2e-3, atol=1e-3 * grad.abs().max().item())


def test_gradient_descent_and_sart_reduce_the_residual():
    from nefct_hip.recon import gradient_descent, sart
    import nefct_hip as nh
    vol, det, ang, truth, b = _setup()
    x, hist = gradient_descent(b, vol, det, ang, steps=8, lam=1e-3, eps=1e-3)
    assert all(h2 <= h1 * (1 + 1e-6) for h1, h2 in zip(hist, hist[1:]))
    assert hist[-1] < 0.5 * hist[0]
    xs = sart(b, vol, det, ang, iters=4)
    r0 = torch.linalg.vector_norm(b).item()
    r1 = torch.linalg.vector_norm(nh.project(xs, vol, det, ang) - b).item()
    assert r1 < 0.3 * r0


def test_conj_grad_approaches_the_dense_normal_equation_solution():
    """Tiny problem: the normal matrix A^T A + mu W^T W is assembled column by
    column from the GPU operators (FP64 on the host) and solved densely; CG
    (conj_grad.py:37-62) reaches that solution."""
    import nefct_hip as nh
    from nefct_hip.recon import conj_grad, tv_diff, tv_diff_adjoint
    vol = nh.VolumeGeometry.cube((6, 6, 5), 2.0)
    det = nh.ConeGeometry.flat(100.0, 150.0, 16, 1.6, 14, 1.6)
    ang = torch.tensor((2 * math.pi * np.arange(20) / 20).astype(np.float32))
    g = torch.Generator().manual_seed(4)
    truth = torch.rand(vol.shape_zyx, generator=g).to(DEV)
    b = nh.project(truth, vol, det, ang) + 0.01 * torch.randn((20, 14, 16), generator=g).to(DEV)
    mu = 0.05
    n = truth.numel()
    M = np.zeros((n, n))
    for j in range(n):
        e = torch.zeros(n, device=DEV)
        e[j] = 1.0
        e = e.reshape(vol.shape_zyx)
        col = nh.backproject(nh.project(e, vol, det, ang), vol, det, ang) + mu * tv_diff_adjoint(tv_diff(e))
        M[:, j] = col.reshape(-1).double().cpu().numpy()
    rhs = nh.backproject(b, vol, det, ang).reshape(-1).double().cpu().numpy()
    want = np.linalg.solve(0.5 * (M + M.T), rhs)
    x = conj_grad(b, vol, det, ang, n_iter=120, mu=mu)
    got = x.reshape(-1).double().cpu().numpy()
    assert np.linalg.norm(got - want) <= 2e-3 * np.linalg.norm(want)


def test_tv_kernel_matches_torch_value_and_gradient():
    from nefct_hip.recon import TVSmooth, tv_smooth_torch
    g = torch.Generator().manual_seed(8)
    x = torch.rand((13, 17, 70), generator=g).to(DEV)
    x[4:8, 3:9, 10:40] += 1.0
    for eps in (1e-3, 0.1):
        xa = x.clone().requires_grad_(True)
        va = TVSmooth.apply(xa, eps)
        va.backward()
        xb = x.clone().double().requires_grad_(True)
        vb = tv_smooth_torch(xb, eps)
        vb.backward()
        assert va.item() == pytest.approx(vb.item(), rel=1e-6)
        assert torch.allclose(xa.grad.double(), xb.grad, rtol=1e-4, atol=1e-5)


def test_tv_kernel_and_descent_step_vs_oracle(oracle):
    """The HIP TV value and gradient against the oracle's restatement of the
    reference differences (total_variation.py:19-35) and their transpose
    (:125-149): TV(x) = sum sqrt(d^2 + eps^2) over d = tv_forward(x), gradient
    tv_adjoint(d / sqrt(d^2 + eps^2)); and the fused descent step
    x - step (g + lam grad TV(x)) against the same, in FP64 on the host."""
    from nefct_hip.recon import TVSmooth, tv_descent_step
    g = torch.Generator().manual_seed(9)
    x = torch.rand((11, 19, 67), generator=g)
    x[3:7, 2:12, 5:50] += 1.0
    xn = x.double().numpy()
    for eps in (1e-3, 0.1):
        d = oracle.tv_forward(xn)
        r = np.sqrt(d * d + eps * eps)
        want_v = r.sum()
        want_g = oracle.tv_adjoint(d / r)
        xa = x.to(DEV).requires_grad_(True)
        va = TVSmooth.apply(xa, eps)
        va.backward()
        assert va.item() == pytest.approx(want_v, rel=1e-6)
        got = xa.grad.double().cpu().numpy()
        assert np.linalg.norm(got - want_g) <= 1e-5 * np.linalg.norm(want_g)
        dg = torch.randn(x.shape, generator=g)
        lam, step = 0.3, 0.05
        out = tv_descent_step(x.to(DEV), dg.to(DEV), eps, lam, step).double().cpu().numpy()
        want = xn - step * (dg.double().numpy() + lam * want_g)
        assert np.linalg.norm(out - want) <= 1e-6 * np.linalg.norm(want)


def test_fused_and_graph_descent_match_autograd_descent():
    from nefct_hip.recon import gradient_descent, power_iteration
    vol, det, ang, truth, b = _setup(24, 20)
    Lp = power_iteration(vol, det, ang, device=DEV)
    xa, ha = gradient_descent(b, vol, det, ang, steps=5, lam=1e-3, eps=1e-2, lipschitz=Lp, fused=False)
    xf, hf = gradient_descent(b, vol, det, ang, steps=5, lam=1e-3, eps=1e-2, lipschitz=Lp, fused=True)
    xg, hg = gradient_descent(b, vol, det, ang, steps=5, lam=1e-3, eps=1e-2, lipschitz=Lp, fused=True, graph=True)
    assert np.allclose(ha, hf, rtol=1e-5)
    assert np.allclose(hf, hg, rtol=1e-6)
    scale = xa.abs().max().item()
    assert torch.allclose(xf, xa, rtol=1e-4, atol=1e-5 * scale)
    assert torch.allclose(xg, xf, rtol=1e-5, atol=1e-6 * scale)
