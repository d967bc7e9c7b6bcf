"""The batched workloads at their real sizes.

C4: ten respiratory phases of one 256 x 256 x 160 base volume, each warped by
its own Gaussian-bump displacement field and seen from 40-120 views of a
384 x 288 detector.  The batched warp -> forward -> adjoint path equals the
per-phase plain calls, the base-volume gradient passes the dot test, the
displacement gradient matches central differences, and one reconstruction
iteration captured in a HIP graph replays like the eager iteration.

C5: a subset of the 64 uneven problems (every volume side, up to 360 views
and 512 x 512 detectors) through the batched pair equals per-problem calls.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rel(a, b):
    return (torch.linalg.norm((a - b).double()) / torch.linalg.norm(b.double())).item()


@pytest.fixture(scope="module")
def c4():
    from nefct_hip import configs
    from nefct_hip.batch import SiddonBatch
    from nefct_hip.phantom import ellipsoid_phantom, gaussian_dvf
    phases = configs.c4()
    vol = phases[0].vol
    base = ellipsoid_phantom(vol, phases[0].phantom_seed, 10, device=DEV)
    dvfs = [gaussian_dvf(vol.shape_zyx, p.dvf_seed, voxel_mm=vol.voxel, device=DEV) for p in phases]
    batch = SiddonBatch([(p.vol, p.det, p.angles) for p in phases], device=DEV)
    return phases, base, dvfs, batch


def test_c4_shapes(c4):
    phases, base, dvfs, batch = c4
    assert len(phases) == 10 and base.shape == (160, 256, 256)
    assert all(d.shape == (3, 160, 256, 256) for d in dvfs)
    assert all(40 <= len(p.angles) <= 120 for p in phases)
    assert all((p.det.na, p.det.nb) == (384, 288) for p in phases)
    # the fields move voxels by up to 10 mm along z (1 mm voxels)
    peak = max(d[2].abs().max().item() for d in dvfs)
    assert 1.0 < peak <= 10.0 * 8


def test_c4_batched_equals_per_phase_plain(c4):
    import nefct_hip as nh
    from nefct_hip.deform import deform
    from nefct_hip.fourd import warp_project
    phases, base, dvfs, batch = c4
    outs = warp_project(batch, [base] * len(phases), dvfs, boundary=1)
    for p, d, o in zip(phases, dvfs, outs):
        ref = nh.project(deform(base, d[0], d[1], d[2], 1), p.vol, p.det, torch.tensor(p.angles))
        assert torch.allclose(o, ref, rtol=1e-6, atol=1e-6 * ref.abs().max().item())


def test_c4_base_gradient_dot_test_and_plain_adjoint(c4):
    import nefct_hip as nh
    from nefct_hip.deform import deform_adjoint_op
    from nefct_hip.fourd import warp_project
    phases, base, dvfs, batch = c4
    g = torch.Generator(device=DEV).manual_seed(41)
    x = torch.randn(base.shape, generator=g, device=DEV).requires_grad_(True)
    ys = [torch.randn(s.shape, generator=g, device=DEV) for s in batch.sinograms]
    outs = warp_project(batch, [x] * len(phases), dvfs, boundary=1)
    loss = sum((o.double() * y.double()).sum() for o, y in zip(outs, ys))
    loss.backward()
    lhs = loss.item()
    rhs = (x.detach().double() * x.grad.double()).sum().item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    # the same gradient from per-phase plain adjoint calls
    want = torch.zeros_like(base)
    for p, d, y in zip(phases, dvfs, ys):
        bp = nh.backproject(y, p.vol, p.det, torch.tensor(p.angles))
        want += deform_adjoint_op(bp, d[0], d[1], d[2], 1)
    assert _rel(x.grad, want) < 1e-5


def _interior_bump(shape_zyx, seed):
    """[3, nz, ny, nx] direction: one Gaussian bump per component (sigma 12
    voxels, positive, centre within the middle third of each axis), largest
    component 1 voxel.  It moves only voxels far from the faces, so no sample
    crosses the warp's validity boundary."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz, ny, nx = shape_zyx
    iz = torch.arange(nz, dtype=torch.float64, device=DEV)[:, None, None]
    iy = torch.arange(ny, dtype=torch.float64, device=DEV)[None, :, None]
    ix = torch.arange(nx, dtype=torch.float64, device=DEV)[None, None, :]
    comps = []
    for _ in range(3):
        cz, cy, cx = (rng.uniform(n / 3, 2 * n / 3) for n in (nz, ny, nx))
        amp = rng.uniform(0.5, 1.0)
        comps.append(amp * torch.exp(-((ix - cx) ** 2 + (iy - cy) ** 2 + (iz - cz) ** 2) / (2 * 12.0 ** 2)))
    d = torch.stack(comps)
    return (d / d.abs().max()).float()


def test_c4_displacement_gradient_central_differences(c4):
    """The displacement-field gradient of loss = sum_p <A_p W(d_p) x, y_p>
    against a difference quotient along a smooth interior direction of one
    phase's field.  The projector is linear, so the difference of the loss
    is taken as <A_p (W(d + eps delta) x - W(d) x), y_p>: only the warped
    volumes are differenced, and the float32 rounding of the 8.8e7
    projection values (which would swamp a 1e-3 change of the summed loss)
    does not enter.  The C4 fields move voxels along z only, so every x and
    y sample sits exactly on a cell boundary, where trilinear interpolation
    has a kink and the warp (like floor) takes the cell on the positive
    side: the quotient is one-sided, along a direction that is positive in
    every component, which is the derivative the kernel computes."""
    import nefct_hip as nh
    from nefct_hip.deform import deform_op
    from nefct_hip.fourd import warp_project
    phases, base, dvfs, batch = c4
    g = torch.Generator(device=DEV).manual_seed(42)
    ys = [torch.randn(s.shape, generator=g, device=DEV) for s in batch.sinograms]
    ds = [d.clone().requires_grad_(True) for d in dvfs]
    outs = warp_project(batch, [base] * len(phases), ds, boundary=1)
    loss = sum((o.double() * y.double()).sum() for o, y in zip(outs, ys))
    loss.backward()
    eps = 1e-3
    for k in (0, 7):
        p = phases[k]
        delta = _interior_bump(p.vol.shape_zyx, 900 + k)
        d0, dp = dvfs[k], dvfs[k] + eps * delta
        diff = deform_op(base, dp[0], dp[1], dp[2], 1) - deform_op(base, d0[0], d0[1], d0[2], 1)
        proj = nh.project(diff, p.vol, p.det, torch.tensor(p.angles))
        fd = (proj.double() * ys[k].double()).sum().item() / eps
        an = (ds[k].grad.double() * delta.double()).sum().item()
        assert abs(an) > 0
        assert abs(fd - an) <= 1e-2 * abs(an), (k, fd, an)


def test_c4_graph_captured_iteration(c4):
    """One gradient-descent iteration on the base volume (warp every phase,
    batched forward, residual, batched adjoint, transposed warps summed into
    the base gradient, update), captured once and replayed three times with
    eager work in between: every replay equals the eager iteration."""
    from nefct_hip.deform import deform_adjoint_op, deform_op
    phases, base, dvfs, batch = c4
    targets = [s.clone() for s in batch.sinograms]
    g = torch.Generator(device=DEV).manual_seed(43)
    for t in targets:
        t.copy_(torch.rand(t.shape, generator=g, device=DEV))
    x = torch.zeros_like(base)
    grad = torch.zeros_like(base)
    step = 1e-5

    def iteration():
        for p, d in enumerate(dvfs):
            batch.volumes[p].copy_(deform_op(x, d[0], d[1], d[2], 1))
        batch.project()
        for s, t in zip(batch.sinograms, targets):
            s.sub_(t)
        batch.backproject()
        grad.zero_()
        for p, d in enumerate(dvfs):
            grad.add_(deform_adjoint_op(batch.volumes[p], d[0], d[1], d[2], 1))
        x.add_(grad, alpha=-step)

    x0 = base * 0.5
    eager = []
    x.copy_(x0)
    for _ in range(3):
        iteration()
        eager.append(x.clone())
    x.copy_(x0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        iteration()          # warm-up outside the graph
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        iteration()
    x.copy_(x0)
    torch.cuda.synchronize()
    for k in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.allclose(x, eager[k], rtol=1e-5, atol=1e-6 * eager[k].abs().max().item()), k
        _ = (x * 2).sum().item()   # eager work between replays


def test_c5_subset_batched_equals_plain():
    """Eight C5 problems at their real sizes (one of each side, the largest
    and smallest view counts), batched forward and adjoint against per-problem
    plain calls."""
    import nefct_hip as nh
    from nefct_hip import configs
    from nefct_hip.batch import SiddonBatch
    from nefct_hip.phantom import ellipsoid_phantom
    probs = configs.c5()
    pick = {}
    for i, p in enumerate(probs):
        pick.setdefault(p.vol.grid[0], i)
    order = sorted(range(len(probs)), key=lambda i: len(probs[i].angles))
    chosen = sorted(set(pick.values()) | {order[0], order[-1], order[len(order) // 2]})
    sub = [probs[i] for i in chosen]
    assert {p.vol.grid[0] for p in sub} == {64, 96, 128, 192, 256}
    b = SiddonBatch([(p.vol, p.det, p.angles) for p in sub], device=DEV)
    for v, p in zip(b.volumes, sub):
        v.copy_(ellipsoid_phantom(p.vol, p.phantom_seed, 10, device=DEV))
    b.project()
    for p, v, s in zip(sub, b.volumes, b.sinograms):
        ref = nh.project(v, p.vol, p.det, torch.tensor(p.angles))
        assert torch.allclose(s, ref, rtol=1e-6, atol=1e-6 * ref.abs().max().item())
    g = torch.Generator(device=DEV).manual_seed(44)
    for s in b.sinograms:
        s.copy_(torch.rand(s.shape, generator=g, device=DEV))
    b.backproject()
    for p, v, s in zip(sub, b.volumes, b.sinograms):
        ref = nh.backproject(s, p.vol, p.det, torch.tensor(p.angles))
        assert _rel(v, ref) < 1e-5
