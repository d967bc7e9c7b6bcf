"""Batched Siddon pair (persistent launches over uneven problems): every
problem's result equals the plain per-problem call, the batched autograd
Function is the exact adjoint, and a captured HIP graph replays correctly."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _problems(seed=0, n=5):
    import nefct_hip as nh
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(n):
        side = int(rng.choice([16, 24, 32, 40]))
        vol = nh.VolumeGeometry.cube((side, side + 4 * (i % 2), side - 4 * (i % 3 == 0)), 1.0 + 0.25 * i)
        sod = float(rng.uniform(60, 120))
        det = nh.ConeGeometry.flat(sod, 1.5 * sod, 2 * side + 3 * i, 1.2, 2 * side - i, 1.1)
        nv = int(rng.integers(5, 17))
        ang = rng.uniform(0, 2 * np.pi, nv).astype(np.float32)
        probs.append((vol, det, ang))
    return probs


def test_batch_matches_plain_calls():
    import nefct_hip as nh
    from nefct_hip.batch import SiddonBatch
    probs = _problems(1, 6)
    b = SiddonBatch(probs)
    g = torch.Generator().manual_seed(2)
    for v in b.volumes:
        v.copy_(torch.rand(v.shape, generator=g))
    b.project()
    for (vol, det, ang), v, s in zip(probs, b.volumes, b.sinograms):
        ref = nh.project(v, vol, det, torch.tensor(ang))
        assert torch.allclose(s, ref, rtol=1e-6, atol=1e-6 * ref.abs().max().item())
    for s in b.sinograms:
        s.copy_(torch.rand(s.shape, generator=g))
    b.backproject()
    for (vol, det, ang), v, s in zip(probs, b.volumes, b.sinograms):
        ref = nh.backproject(s, vol, det, torch.tensor(ang))
        assert torch.allclose(v, ref, rtol=1e-5, atol=1e-5 * ref.abs().max().item())


def test_batched_autograd_is_adjoint():
    from nefct_hip.batch import SiddonBatch, batched_project
    probs = _problems(3, 4)
    b = SiddonBatch(probs)
    g = torch.Generator().manual_seed(4)
    xs = [torch.randn(v.shape, generator=g).to(DEV).requires_grad_(True) for v in b.volumes]
    ys = [torch.randn(s.shape, generator=g).to(DEV) for s in b.sinograms]
    outs = batched_project(b, xs)
    loss = sum((o.double() * y.double()).sum() for o, y in zip(outs, ys))
    loss.backward()
    lhs = loss.item()
    rhs = sum((x.detach().double() * x.grad.double()).sum().item() for x in xs)
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_graph_captured_iteration_replays():
    """A gradient step (project, residual, backproject, update) captured in a
    HIP graph: every replay matches the eager step."""
    from nefct_hip.batch import SiddonBatch
    probs = _problems(5, 3)
    b = SiddonBatch(probs)
    g = torch.Generator().manual_seed(6)
    targets = [torch.rand(s.shape, generator=g).to(DEV) for s in b.sinograms]
    x0 = [torch.rand(v.shape, generator=g).to(DEV) * 0.01 for v in b.volumes]

    def step():
        b.project()
        for s, t in zip(b.sinograms, targets):
            s.sub_(t)
        b.backproject()

    # eager reference sequence
    eager = []
    for v, x in zip(b.volumes, x0):
        v.copy_(x)
    for _ in range(3):
        step()
        for v in b.volumes:
            v.mul_(-1e-4).add_(v)   # deterministic in-place update
        eager.append([v.clone() for v in b.volumes])
    # capture
    for v, x in zip(b.volumes, x0):
        v.copy_(x)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()   # warm-up outside the graph
    torch.cuda.current_stream().wait_stream(s)
    for v, x in zip(b.volumes, x0):
        v.copy_(x)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
        for v in b.volumes:
            v.mul_(-1e-4).add_(v)
    for v, x in zip(b.volumes, x0):
        v.copy_(x)
    torch.cuda.synchronize()
    for k in range(3):
        graph.replay()
        torch.cuda.synchronize()
        for v, e in zip(b.volumes, eager[k]):
            assert torch.allclose(v, e, rtol=1e-5, atol=1e-6 * e.abs().max().item()), k
        # eager work between replays
        _ = (b.volumes[0] * 2).sum().item()


def test_warp_project_batch_gradients():
    """4-D path: shared base volume, per-phase warps and views.  Gradients to
    the base volume equal the sum of the per-phase transposes; gradients to the
    displacement fields match central differences inside a trilinear cell."""
    import nefct_hip as nh
    from nefct_hip.batch import SiddonBatch
    from nefct_hip.fourd import warp_project
    from nefct_hip.deform import deform, deform_adjoint_op
    vol = nh.VolumeGeometry.cube((24, 24, 16), 1.5)
    det = nh.ConeGeometry.flat(80.0, 120.0, 48, 1.1, 32, 1.1)
    rng = np.random.default_rng(7)
    phases = [(vol, det, rng.uniform(0, 2 * np.pi, int(rng.integers(4, 9))).astype(np.float32)) for _ in range(3)]
    b = SiddonBatch(phases)
    g = torch.Generator().manual_seed(8)
    base = torch.rand(vol.shape_zyx, generator=g).to(DEV).requires_grad_(True)
    dvfs = [((torch.floor(torch.rand((3,) + vol.shape_zyx, generator=g) * 4 - 2)
              + 0.2 + 0.6 * torch.rand((3,) + vol.shape_zyx, generator=g)).to(DEV).requires_grad_(True))
            for _ in phases]
    ys = [torch.randn(s.shape, generator=g).to(DEV) for s in b.sinograms]
    outs = warp_project(b, [base] * 3, dvfs, boundary=1)
    loss = sum((o.double() * y.double()).sum() for o, y in zip(outs, ys))
    loss.backward()
    # base gradient = sum_p W_p^T A_p^T y_p
    want = torch.zeros_like(base)
    for (v, d_, a), y, d in zip(phases, ys, dvfs):
        bp = nh.backproject(y, v, d_, torch.tensor(a))
        want += deform_adjoint_op(bp, d.detach()[0], d.detach()[1], d.detach()[2], 1)
    assert torch.allclose(base.grad, want, rtol=1e-5, atol=1e-5 * want.abs().max().item())
    # displacement gradient by central differences (exact inside a cell)
    eps = 0.1
    k = 1
    e = torch.zeros_like(dvfs[k])
    e[2] = eps
    def f(dv):
        w = deform(base.detach(), dv[0], dv[1], dv[2], 1)
        return (nh.project(w, phases[k][0], phases[k][1], torch.tensor(phases[k][2])).double() * ys[k].double()).sum()
    fd = (f(dvfs[k].detach() + e) - f(dvfs[k].detach() - e)) / (2 * eps)
    an = (dvfs[k].grad[2].double()).sum() * 0 + (dvfs[k].grad[2].double() * (e[2] / eps)).sum()
    assert abs(fd.item() - an.item()) <= 2e-3 * abs(an.item()) + 1e-6


@pytest.mark.parametrize("signed", [False, True])
def test_batched_adjoint_lane_stride_bitwise(signed):
    """Neighbouring lanes two detector rows apart (NEFCT_F_LANE_STRIDE2) in
    the batched adjoint: the fixed-point sums are order-independent, so the
    volumes are bitwise those of lane stride 1; tune() keeps one of the two."""
    from nefct_hip import _lib
    from nefct_hip.batch import SiddonBatch
    probs = _problems(5, 5)
    b = SiddonBatch(probs)
    g = torch.Generator().manual_seed(6)
    for s in b.sinograms:
        s.copy_(torch.randn(s.shape, generator=g) if signed else torch.rand(s.shape, generator=g))
    b.backproject()
    one = [v.clone() for v in b.volumes]
    b.flags |= _lib.F_LANE_STRIDE2
    b.backproject()
    assert all(torch.equal(a, v) for a, v in zip(one, b.volumes))
    rec = b.tune()
    assert rec["adjoint"] in ("stride1", "stride2") and set(rec["ms"]) == {"stride1", "stride2"}
    b.backproject()
    assert all(torch.equal(a, v) for a, v in zip(one, b.volumes))


def test_autograd_launch_shape_flags():
    """forward_project / back_project take launch-shape flags (the pair's
    result is unchanged: bitwise for the fixed-point adjoint) and refuse the
    reference rule, which is not an adjoint pair."""
    import nefct_hip as nh
    from nefct_hip import _lib
    vol, det, ang = _problems(7, 1)[0]
    a = torch.tensor(ang)
    g = torch.Generator().manual_seed(8)
    x0 = torch.rand(vol.shape_zyx, generator=g).to(DEV)
    y = torch.randn((len(ang), det.nb, det.na), generator=g).to(DEV)
    grads = []
    for fl in (0, _lib.F_LANE_STRIDE2 | _lib.F_TILE_32x16x8, _lib.F_FWD_TILE_32x8 | _lib.fwd_vg(3)):
        x = x0.clone().requires_grad_(True)
        (nh.forward_project(x, vol, det, a, flags=fl) * y).sum().backward()
        grads.append(x.grad)
    assert torch.equal(grads[0], grads[1]) and torch.equal(grads[0], grads[2])
    with pytest.raises(_lib.NefctError):
        nh.forward_project(x0, vol, det, a, flags=_lib.REF_FORWARD)
