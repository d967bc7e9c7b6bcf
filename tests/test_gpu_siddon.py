"""GPU parity of the Siddon pair against the CPU oracle (HIP path through the C ABI).

Bars: exact forward within 1e-5 relative L2 of the FP64 oracle; reference-mode
forward/adjoint within 1e-5 of the FP32 restatement of
nefct/cc/siddon_3d/siddon.cu.cc; exact adjoint within 1e-5 of the FP64
transpose; GPU dot test <Ax, y> = <x, A^T y> to 1e-6 with FP64 accumulation.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


def _nh():
    import nefct_hip
    nefct_hip.load()
    return nefct_hip


def _relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(oracle, vol, det, nv, seed=0, views=None, k=8):
    nh = _nh()
    from nefct_hip.phantom import ellipsoid_phantom
    ang = (2 * math.pi * np.arange(nv) / nv).astype(np.float32) + np.float32(0.013)
    if views is not None:
        ang = ang[views]
    img = ellipsoid_phantom(vol, seed, k)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(det.cyl, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    return nh, ang, img, ov, od


def _geoms():
    from nefct_hip import ConeGeometry, VolumeGeometry
    return [
        ("c1-like", VolumeGeometry.cube((64, 64, 64), 2.0), ConeGeometry.flat(1000.0, 1500.0, 128, 3.0, 128, 3.0)),
        ("aniso-offset", VolumeGeometry((48, 40, 36), (3.0, -2.0, 1.5), (48 * 1.1, 40 * 1.1, 36 * 0.9)),
         ConeGeometry.flat(300.0, 500.0, 70, 2.3, 45, 2.1, ai=0.3, bi=0.6)),
        ("cyli", VolumeGeometry.cube((48, 48, 40), 1.5),
         ConeGeometry("cyli", 400.0, 700.0, 90, 0.0035, 0.5, 50, 1.9, 0.5)),
    ]


@pytest.mark.parametrize("gi", [0, 1, 2])
def test_forward_exact_vs_fp64_oracle(oracle, gi):
    name, vol, det = _geoms()[gi]
    nh, ang, img, ov, od = _case(oracle, vol, det, 36, seed=gi + 1)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    ref = oracle.exact_fwd(ov, od, ang.astype(np.float64), img.double().numpy())
    assert got.shape == ref.shape
    assert _relerr(got, ref) < 1e-5, (name, _relerr(got, ref))


@pytest.mark.parametrize("gi,nv,seed", [(0, 36, 7), (2, 36, 9), (0, 24, 3)])
def test_forward_reference_mode_vs_restatement(oracle, gi, nv, seed):
    """(0, 24, 3) holds rays that miss the volume box by a hair at its bottom
    corners, which the reference's slab attribution still credits to an
    in-range voxel (view 3 row 30 column 88 and three more)."""
    name, vol, det = _geoms()[gi]
    nh, ang, img, ov, od = _case(oracle, vol, det, nv, seed=seed)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang), flags=nh.REF_FORWARD).cpu().numpy()
    ref = oracle.ref_fwd64(ov, od, ang.astype(np.float64), img.double().numpy(), flags=oracle.REF_FWD_FLAGS)
    assert _relerr(got, ref) < 1e-5, (name, _relerr(got, ref))
    # and within the FP32 restatement's own rounding noise of it
    ref32 = oracle.ref_fwd(ov, od, ang, img.numpy(), flags=oracle.REF_FWD_FLAGS)
    assert _relerr(got, ref32) < 3e-5, (name, _relerr(got, ref32))


@pytest.mark.parametrize("gi", [0, 1, 2])
def test_adjoint_exact_vs_fp64_transpose(oracle, gi):
    name, vol, det = _geoms()[gi]
    nh, ang, img, ov, od = _case(oracle, vol, det, 24, seed=gi + 3)
    rng = np.random.default_rng(gi)
    y = rng.uniform(0, 1, size=(24, det.nb, det.na)).astype(np.float32)
    got = nh.backproject(torch.tensor(y, device=DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    ref = oracle.exact_bp(ov, od, ang.astype(np.float64), y.astype(np.float64))
    assert got.shape == ref.shape
    assert _relerr(got, ref) < 1e-5, (name, _relerr(got, ref))


@pytest.mark.parametrize("gi", [0, 1, 2])
def test_adjoint_reference_mode_vs_restatement(oracle, gi):
    name, vol, det = _geoms()[gi]
    nh, ang, img, ov, od = _case(oracle, vol, det, 24, seed=gi + 5)
    rng = np.random.default_rng(gi + 10)
    y = rng.uniform(0, 1, size=(24, det.nb, det.na)).astype(np.float32)
    got = nh.backproject(torch.tensor(y, device=DEV), vol, det, torch.tensor(ang), flags=nh.REF_ADJOINT).cpu().numpy()
    ref = oracle.ref_bp64(ov, od, ang.astype(np.float64), y.astype(np.float64), flags=oracle.REF_BP_FLAGS)
    assert _relerr(got, ref) < 1e-5, (name, _relerr(got, ref))
    ref32 = oracle.ref_bp(ov, od, ang, y, flags=oracle.REF_BP_FLAGS)
    assert _relerr(got, ref32) < 3e-5, (name, _relerr(got, ref32))


@pytest.mark.parametrize("gi", [0, 1, 2])
def test_dot_test_gpu(gi):
    nh = _nh()
    name, vol, det = _geoms()[gi]
    g = torch.Generator(device="cpu").manual_seed(gi)
    nv = 20
    ang = torch.rand(nv, generator=g) * 2 * math.pi
    x = torch.randn(vol.shape_zyx, generator=g).to(DEV)
    y = torch.randn((nv, det.nb, det.na), generator=g).to(DEV)
    Ax = nh.project(x, vol, det, ang)
    ATy = nh.backproject(y, vol, det, ang)
    lhs = torch.sum(Ax.double() * y.double()).item()
    rhs = torch.sum(x.double() * ATy.double()).item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)


def test_count_vs_oracle(oracle):
    name, vol, det = _geoms()[0]
    nh, ang, img, ov, od = _case(oracle, vol, det, 12)
    got = nh.count_intersections(vol, det, torch.tensor(ang)).cpu().numpy()
    total, per = oracle.exact_count(ov, od, ang.astype(np.float64), per_ray=True)
    assert abs(int(got.sum()) - total) <= 1e-5 * total
    assert np.mean(got != per) < 1e-3


def test_steep_rays_z_dominant(oracle):
    """Detector taller than wide at short distance: many rays step in z."""
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry.cube((24, 24, 24), 1.0)
    det = ConeGeometry.flat(30.0, 45.0, 12, 4.0, 40, 6.0)
    nh, ang, img, ov, od = _case(oracle, vol, det, 6, k=5)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    ref = oracle.exact_fwd(ov, od, ang.astype(np.float64), img.double().numpy())
    assert _relerr(got, ref) < 1e-5
    y = np.random.default_rng(1).uniform(size=ref.shape).astype(np.float32)
    gb = nh.backproject(torch.tensor(y, device=DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    rb = oracle.exact_bp(ov, od, ang.astype(np.float64), y.astype(np.float64))
    assert _relerr(gb, rb) < 1e-5


def test_edge_cases_missing_rays_and_ragged_detector(oracle):
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry.cube((16, 16, 16), 1.0)
    # detector much wider than the volume shadow: most rays miss; na, nb not tile multiples
    det = ConeGeometry.flat(100.0, 150.0, 67, 3.0, 5, 3.0)
    nh, ang, img, ov, od = _case(oracle, vol, det, 3, k=4)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    ref = oracle.exact_fwd(ov, od, ang.astype(np.float64), img.double().numpy())
    assert np.all(got[ref == 0] == 0)
    assert _relerr(got, ref) < 1e-5
    # single view, single row
    det1 = ConeGeometry.flat(100.0, 150.0, 9, 1.0, 1, 1.0)
    got1 = nh.project(img.to(DEV), vol, det1, torch.tensor([0.3])).cpu().numpy()
    od1 = oracle.det_struct(0, 100.0, 150.0, 9, 1.0, 0.5, 1, 1.0, 0.5)
    ref1 = oracle.exact_fwd(ov, od1, [np.float64(np.float32(0.3))], img.double().numpy())
    assert _relerr(got1, ref1) < 1e-5


def test_bad_arguments_raise():
    nh = _nh()
    from nefct_hip import ConeGeometry, VolumeGeometry
    vol = VolumeGeometry.cube((8, 8, 8), 1.0)
    det = ConeGeometry.flat(100.0, 150.0, 4, 1.0, 4, 1.0)
    x = torch.zeros(vol.shape_zyx, device=DEV)
    with pytest.raises(nh.NefctError):
        nh.project(x, vol, ConeGeometry.flat(100.0, 150.0, 4, float("nan"), 4, 1.0), [0.0])
    with pytest.raises(nh.NefctError):
        nh.project(x[:4], vol, det, [0.0])
    with pytest.raises(nh.NefctError):
        nh.project(x.cpu(), vol, det, [0.0])
    aniso = VolumeGeometry((8, 8, 8), (0, 0, 0), (8.0, 16.0, 8.0))
    with pytest.raises(nh.NefctError):
        nh.project(x, aniso, det, [0.0], flags=nh.REF_FORWARD)


def test_full_size_c2_subset_properties():
    """At the C2 volume/detector size on 6 views: project(1) equals the ray's
    chord with the box (computed in FP64 by the slab method), and the dot test."""
    nh = _nh()
    from nefct_hip.configs import c2
    wl = c2()
    vol, det = wl.vol, wl.det
    ang = torch.tensor(wl.angles()[::120].copy())
    ones = torch.ones(vol.shape_zyx, device=DEV)
    p = nh.project(ones, vol, det, ang).double()
    # chords in FP64
    nv = ang.numel()
    a = (det.ai + torch.arange(det.na, dtype=torch.float64, device=DEV)) * det.da - det.na * det.da / 2
    b = (det.bi + torch.arange(det.nb, dtype=torch.float64, device=DEV)) * det.db - det.nb * det.db / 2
    t = ang.double().to(DEV)
    c, s = torch.cos(t)[:, None, None], torch.sin(t)[:, None, None]
    B, A = torch.meshgrid(b, a, indexing="ij")
    A, B = A[None], B[None]
    S = torch.stack([-det.SID * c + 0 * A, -det.SID * s + 0 * A, 0 * A + 0 * c])
    D = torch.stack([(det.SAD - det.SID) * c - A * s, (det.SAD - det.SID) * s + A * c, B + 0 * c])
    d = D - S
    lo = torch.tensor([-s_ / 2 for s_ in vol.size], dtype=torch.float64, device=DEV)
    tin = torch.full_like(A + 0 * c, -1e300)
    tout = torch.full_like(A + 0 * c, 1e300)
    for i in range(3):
        t0 = (lo[i] - S[i]) / d[i]
        t1 = (-lo[i] - S[i]) / d[i]
        tin = torch.maximum(tin, torch.minimum(t0, t1))
        tout = torch.minimum(tout, torch.maximum(t0, t1))
    chord = torch.clamp(tout - tin, min=0) * torch.sqrt((d * d).sum(0))
    assert (torch.norm(p - chord) / torch.norm(chord)).item() < 1e-5
    g = torch.Generator(device="cpu").manual_seed(3)
    x = torch.randn(vol.shape_zyx, generator=g).to(DEV)
    y = torch.randn((nv, det.nb, det.na), generator=g).to(DEV)
    lhs = torch.sum(nh.project(x, vol, det, ang).double() * y.double()).item()
    rhs = torch.sum(x.double() * nh.backproject(y, vol, det, ang).double()).item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_autograd_matches_adjoint():
    nh = _nh()
    name, vol, det = _geoms()[0]
    ang = torch.linspace(0, 3.0, 10)
    x = torch.rand(vol.shape_zyx, device=DEV, requires_grad=True)
    y = torch.rand((10, det.nb, det.na), device=DEV)
    loss = (nh.forward_project(x, vol, det, ang) * y).sum()
    loss.backward()
    ref = nh.backproject(y, vol, det, ang)
    assert torch.allclose(x.grad, ref, rtol=1e-5, atol=1e-6 * ref.abs().max().item())


@pytest.mark.parametrize("algo", ["tile", "float", "deterministic", "deterministic-float", "window",
                                  "tile32x16x8-stride2", "tile64x32x8", "tile32x32x16", "tile16x16x32",
                                  "tile64x16x16-stride2", "tile64x32x8-float"])
def test_adjoint_algorithms_agree_with_fp64_transpose(oracle, algo):
    """Every adjoint variant against the FP64 transpose (fixed-point and
    float LDS accumulation, every block shape, lane spacing 1 and 2, the
    slab-window kernel)."""
    nh = _nh()
    from nefct_hip import _lib
    fl = {"tile": 0, "float": _lib.F_ADJ_FLOAT_ATOMIC, "deterministic": _lib.F_DETERMINISTIC,
          "deterministic-float": _lib.F_DETERMINISTIC | _lib.F_ADJ_FLOAT_ATOMIC,
          "window": _lib.F_ADJ_WINDOW,
          "tile32x16x8-stride2": _lib.F_TILE_32x16x8 | _lib.F_LANE_STRIDE2,
          "tile64x32x8": _lib.F_TILE_64x32x8, "tile32x32x16": _lib.F_TILE_32x32x16,
          "tile16x16x32": _lib.F_TILE_16x16x32,
          "tile64x16x16-stride2": _lib.F_TILE_64x16x16 | _lib.F_LANE_STRIDE2,
          "tile64x32x8-float": _lib.F_TILE_64x32x8 | _lib.F_ADJ_FLOAT_ATOMIC}[algo]
    for gi in (0, 1, 2):
        name, vol, det = _geoms()[gi]
        _, ang, img, ov, od = _case(oracle, vol, det, 16, seed=gi + 21)
        y = np.random.default_rng(gi + 30).uniform(0, 1, size=(16, det.nb, det.na)).astype(np.float32)
        got = nh.backproject(torch.tensor(y, device=DEV), vol, det, torch.tensor(ang), flags=fl).cpu().numpy()
        ref = oracle.exact_bp(ov, od, ang.astype(np.float64), y.astype(np.float64))
        assert _relerr(got, ref) < 1e-5, (algo, name, _relerr(got, ref))
        got_r = nh.backproject(torch.tensor(y, device=DEV), vol, det, torch.tensor(ang),
                               flags=fl | nh.REF_ADJOINT).cpu().numpy()
        ref_r = oracle.ref_bp64(ov, od, ang.astype(np.float64), y.astype(np.float64), flags=oracle.REF_BP_FLAGS)
        assert _relerr(got_r, ref_r) < 1e-5, (algo, name, "reference mode", _relerr(got_r, ref_r))


def test_adjoint_many_views(oracle):
    """1100 views: the tile adjoint fills its per-view LDS tables (detector
    shadows, sources) 128 views at a time, so this crosses eight chunk
    boundaries (the last chunk partial).  Against the FP64 transpose (non-negative and mixed-sign
    data, which take different deposit paths), the dot test, and the batched
    adjoint of the same problem."""
    nh = _nh()
    from nefct_hip import ConeGeometry, VolumeGeometry
    from nefct_hip.batch import SiddonBatch
    vol = VolumeGeometry.cube((20, 22, 18), 1.5)
    det = ConeGeometry.flat(80.0, 120.0, 40, 1.3, 30, 1.2)
    nv = 1100
    ang = (2 * math.pi * np.arange(nv) / nv + 0.02).astype(np.float32)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(det.cyl, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    rng = np.random.default_rng(41)
    for y in (rng.uniform(0, 1, size=(nv, det.nb, det.na)), rng.standard_normal((nv, det.nb, det.na))):
        y = y.astype(np.float32)
        got = nh.backproject(torch.tensor(y, device=DEV), vol, det, torch.tensor(ang)).cpu().numpy()
        ref = oracle.exact_bp(ov, od, ang.astype(np.float64), y.astype(np.float64))
        assert _relerr(got, ref) < 1e-5, _relerr(got, ref)
    # seeded: with the global generator the draw depended on the tests run before
    x = torch.randn(vol.shape_zyx, generator=torch.Generator().manual_seed(13)).to(DEV)
    yt = torch.randn((nv, det.nb, det.na), generator=torch.Generator().manual_seed(14)).to(DEV)
    lhs = torch.sum(nh.project(x, vol, det, torch.tensor(ang)).double() * yt.double()).item()
    rhs = torch.sum(x.double() * nh.backproject(yt, vol, det, torch.tensor(ang)).double()).item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)
    b = SiddonBatch([(vol, det, ang)])
    b.sinograms[0].copy_(yt)
    b.backproject()
    ref = nh.backproject(yt, vol, det, torch.tensor(ang))
    assert torch.allclose(b.volumes[0], ref, rtol=1e-5, atol=1e-5 * ref.abs().max().item())


@pytest.mark.parametrize("shift", [0.0, 1.0 / 256])
def test_boundary_aligned_rays(oracle, shift):
    """Rays on (or within 1/256 voxel of) voxel faces: views at multiples of
    45 degrees (the float32 angles leave transverse slopes of 0 or ~1e-8) and
    detector cells whose rays pass through the isocentre planes x = 0, y = 0,
    z = 0 (even sizes, no offsets, one voxel per cell at the isocentre), with
    the volume shifted by 'shift' voxels along each axis.  Stresses the zero-
    and near-zero-slope paths, the block overlap margin and the crossing
    fraction at cell boundaries.  Checks: the transpose (dot test, 1e-6) and
    the agreement of every adjoint kernel with the default one (same weights,
    1e-6) always; the FP64 oracle to 1e-5 when shifted.  Unshifted, a ray lying
    in a face has no unique exact answer (its length belongs to either
    neighbour, decided by rounding), so there the oracle only bounds the
    result loosely (1e-2)."""
    nh = _nh()
    from nefct_hip import ConeGeometry, VolumeGeometry, _lib
    vol = VolumeGeometry((32, 32, 32), (shift, shift, shift), (32.0, 32.0, 32.0))
    det = ConeGeometry.flat(100.0, 200.0, 80, 2.0, 80, 2.0, ai=0.0, bi=0.0)
    ang = (np.arange(8) * (math.pi / 4)).astype(np.float32)
    from nefct_hip.phantom import ellipsoid_phantom
    img = ellipsoid_phantom(vol, 7, 6)
    y = np.random.default_rng(3).uniform(0, 1, size=(ang.size, det.nb, det.na)).astype(np.float32)
    yt = torch.tensor(y, device=DEV)
    base = nh.backproject(yt, vol, det, torch.tensor(ang)).cpu().numpy()
    for fl in (_lib.F_TILE_64x32x8, _lib.F_TILE_16x16x32 | _lib.F_LANE_STRIDE2, _lib.F_ADJ_FLOAT_ATOMIC,
               _lib.F_ADJ_WINDOW):
        bp = nh.backproject(yt, vol, det, torch.tensor(ang), flags=fl).cpu().numpy()
        assert _relerr(bp, base) < 1e-6, (fl, _relerr(bp, base))
    x = torch.randn(vol.shape_zyx, generator=torch.Generator().manual_seed(5)).to(DEV)
    yr = torch.randn((ang.size, det.nb, det.na), generator=torch.Generator().manual_seed(6)).to(DEV)
    lhs = torch.sum(nh.project(x, vol, det, torch.tensor(ang)).double() * yr.double()).item()
    rhs = torch.sum(x.double() * nh.backproject(yr, vol, det, torch.tensor(ang)).double()).item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(det.cyl, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    got = nh.project(img.to(DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    ref = oracle.exact_fwd(ov, od, ang.astype(np.float64), img.double().numpy())
    ref_bp = oracle.exact_bp(ov, od, ang.astype(np.float64), y.astype(np.float64))
    tol = 1e-5 if shift else 1e-2
    assert _relerr(got, ref) < tol, _relerr(got, ref)
    assert _relerr(base, ref_bp) < tol, _relerr(base, ref_bp)


def _bp_into_nan(sino, vol, det, ang, flags):
    """The flat-detector adjoint through the C ABI into a NaN-filled image, so
    a voxel the kernel leaves unwritten stays NaN (a fresh tensor can hold the
    bits of an earlier call's image)."""
    from nefct_hip import _lib, ops
    L = _lib.load()
    nx, ny, nz = vol.grid
    out = torch.full((nz, ny, nx), float("nan"), device=sino.device)
    a = ang.to(device=sino.device, dtype=torch.float32).contiguous()
    ws, nbytes = ops._workspace(list(vol.grid), a.numel(), sino.device)
    st = L.nefct_backproject_flat(ops._ptr(sino.contiguous()), _lib.int3(vol.grid), _lib.float3(vol.center),
                                  _lib.float3(vol.size), ops._ptr(a), a.numel(), det.SID, det.SAD, det.na, det.da,
                                  det.ai, det.nb, det.db, det.bi, ops._ptr(out), int(flags), ops._ptr(ws), nbytes,
                                  ops._vp(ops._stream_ptr(sino.device)))
    _lib.check(st, "backproject_flat")
    torch.cuda.synchronize()
    assert not torch.isnan(out).any(), ("unwritten voxels", hex(flags))
    return out


@pytest.mark.parametrize("signed", [False, True])
def test_fixed_point_adjoint_bitwise_across_block_shapes(oracle, signed):
    """The fixed-point adjoint deposits the same integer pieces per (ray,
    slab) whatever block owns the voxel, so every block shape gives the same
    bits.  Blocks that fit a one-cell halo in LDS (32x32x16, 32x32x8, 32x16x8,
    16x16x32) walk wavefronts of rays that stay within it without range
    tests; 64x32x8 and 64x16x16 do not fit one and always test.  A volume
    that is no multiple of any block side (ragged blocks), views near 45
    degrees (transverse slopes near 1, where a ray can reach two cells past a
    block and takes the tested walk) and a wide detector."""
    nh = _nh()
    from nefct_hip import ConeGeometry, VolumeGeometry, _lib
    vol = VolumeGeometry((70, 45, 37), (2.0, -3.0, 1.0), (70 * 1.0, 45 * 1.0, 37 * 1.2))
    det = ConeGeometry.flat(150.0, 260.0, 120, 1.6, 70, 1.5, ai=0.2, bi=0.4)
    ang = np.concatenate([np.float32(math.pi / 4) + np.float32(0.004) * np.arange(-3, 4),
                          (2 * math.pi * np.arange(9) / 9 + 0.1)]).astype(np.float32)
    rng = np.random.default_rng(17 + int(signed))
    y = (rng.standard_normal if signed else rng.uniform)(size=(ang.size, det.nb, det.na)).astype(np.float32)
    yt = torch.tensor(y, device=DEV)
    angt = torch.tensor(ang)
    base = _bp_into_nan(yt, vol, det, angt, _lib.F_TILE_64x32x8).cpu().numpy()
    for fl in (_lib.F_TILE_32x32x16, _lib.F_TILE_32x32x16 | _lib.F_LANE_STRIDE2, _lib.F_TILE_32x32x8,
               _lib.F_TILE_32x16x8, _lib.F_TILE_16x16x32, _lib.F_TILE_64x16x16):
        got = _bp_into_nan(yt, vol, det, angt, fl).cpu().numpy()
        assert np.array_equal(got.view(np.int32), base.view(np.int32)), (hex(fl), _relerr(got, base))
    ov = oracle.vol_struct(vol.grid, vol.center, vol.size)
    od = oracle.det_struct(det.cyl, det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db, det.bi)
    ref = oracle.exact_bp(ov, od, ang.astype(np.float64), y.astype(np.float64))
    assert _relerr(base, ref) < 1e-5, _relerr(base, ref)


def test_deterministic_adjoint_is_bitwise_reproducible():
    nh = _nh()
    from nefct_hip import _lib
    name, vol, det = _geoms()[0]
    ang = torch.linspace(0, 6.0, 30)
    y = torch.rand((30, det.nb, det.na), device=DEV)
    a = nh.backproject(y, vol, det, ang, flags=_lib.F_DETERMINISTIC)
    b = nh.backproject(y, vol, det, ang, flags=_lib.F_DETERMINISTIC)
    assert torch.equal(a, b)


def test_fixed_point_adjoint_range_and_special_values(oracle):
    """The default adjoint sums fixed-point integers scaled from max |y|:
    bitwise reproducible, exact under power-of-two scaling of y, within 1e-5
    of the FP64 transpose for y spanning 10 decades, zero for y = 0, and NaN
    everywhere for a non-finite y."""
    nh = _nh()
    name, vol, det = _geoms()[1]
    _, ang, img, ov, od = _case(oracle, vol, det, 12, seed=41)
    rng = np.random.default_rng(42)
    y = rng.uniform(-1, 1, size=(12, det.nb, det.na)).astype(np.float32)
    yt = torch.tensor(y, device=DEV)
    a = nh.backproject(yt, vol, det, torch.tensor(ang))
    b = nh.backproject(yt, vol, det, torch.tensor(ang))
    assert torch.equal(a, b)
    big = nh.backproject(yt * 2.0 ** 60, vol, det, torch.tensor(ang))
    small = nh.backproject(yt * 2.0 ** -60, vol, det, torch.tensor(ang))
    assert torch.equal(big, a * 2.0 ** 60) and torch.equal(small, a * 2.0 ** -60)
    y2 = y * (10.0 ** rng.uniform(-10, 0, size=y.shape)).astype(np.float32)
    got = nh.backproject(torch.tensor(y2, device=DEV), vol, det, torch.tensor(ang)).cpu().numpy()
    ref = oracle.exact_bp(ov, od, ang.astype(np.float64), y2.astype(np.float64))
    assert _relerr(got, ref) < 1e-5
    assert nh.backproject(torch.zeros_like(yt), vol, det, torch.tensor(ang)).abs().max().item() == 0.0
    yn = yt.clone()
    yn[3, 5, 7] = float("nan")
    assert torch.isnan(nh.backproject(yn, vol, det, torch.tensor(ang))).all()
