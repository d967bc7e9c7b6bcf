"""GPU parity of the texture-rule warps (clamp-to-edge trilinear warp, inverse
displacement field, shift, rotation; nefct/cc/deform/deform_tex.cu.cc) against
the FP64 oracle, through the C ABI.

Bars: warps within 1e-6 relative L2 of the oracle (coordinates are formed in
FP32 on both sides, weights exact); the inverted field within 2e-5 max abs
(30 FP32 iterations against FP64 ones), and its warp within 1e-5 relative L2.
Edge cases: ragged and odd shapes, samples far outside the volume (clamped),
a non-finite displacement at one voxel, shape and size errors.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _t(a):
    return torch.tensor(np.ascontiguousarray(a), device=DEV)


def _smooth_field(shape, seed, amp):
    """Sum of a few Gaussian bumps per component: smooth, |grad| < 0.3."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = []
    for _ in range(3):
        f = np.zeros(shape)
        for _ in range(4):
            c = rng.uniform([0, 0, 0], [nx, ny, nz])
            s = rng.uniform(5.0, 8.0)
            f += rng.uniform(-amp, amp) * np.exp(-((ix - c[0]) ** 2 + (iy - c[1]) ** 2 + (iz - c[2]) ** 2) / (2 * s * s))
        out.append(f.astype(np.float32))
    return out


SHAPES = [(17, 23, 70), (9, 31, 33)]


@pytest.mark.parametrize("shape", SHAPES)
def test_deform_tex_vs_oracle(oracle, shape):
    from nefct_hip.deform import deform_tex
    rng = np.random.default_rng(21)
    img = rng.uniform(size=shape).astype(np.float32)
    # displacements of up to 6 voxels push samples past every face (clamped)
    mx, my, mz = (rng.uniform(-6, 6, size=shape).astype(np.float32) for _ in range(3))
    out = deform_tex(_t(img), _t(mx), _t(my), _t(mz)).cpu().numpy()
    assert _rel(out, oracle.deform_tex(img, mx, my, mz)) < 1e-6


@pytest.mark.parametrize("shape", SHAPES)
def test_invert_dvf_and_deform_invert_vs_oracle(oracle, shape):
    from nefct_hip.deform import deform_invert_tex, invert_dvf
    m = _smooth_field(shape, 22, amp=2.0)
    inv = invert_dvf(_t(m[0]), _t(m[1]), _t(m[2])).cpu().numpy()
    want = oracle.invert_dvf(*m, iters=30)
    for k in range(3):
        assert np.max(np.abs(inv[k] - want[k])) < 2e-5
    img = np.random.default_rng(23).uniform(size=shape).astype(np.float32)
    out = deform_invert_tex(_t(img), _t(m[0]), _t(m[1]), _t(m[2])).cpu().numpy()
    assert _rel(out, oracle.deform_invert_tex(img, *m)) < 1e-5


def test_invert_dvf_iterations_argument(oracle):
    from nefct_hip.deform import invert_dvf
    shape = (8, 10, 12)
    m = _smooth_field(shape, 24, amp=1.5)
    for iters in (0, 1, 3):
        inv = invert_dvf(_t(m[0]), _t(m[1]), _t(m[2]), iters).cpu().numpy()
        want = oracle.invert_dvf(*m, iters=iters)
        for k in range(3):
            assert np.max(np.abs(inv[k] - want[k])) < 1e-5


@pytest.mark.parametrize("s", [(1.25, -2.5, 0.75), (0.0, 0.0, -3.5), (40.0, -40.0, 2.0)])
def test_shift_vs_oracle(oracle, s):
    from nefct_hip.deform import shift_tex, shift_z_tex
    shape = (11, 13, 67)
    img = np.random.default_rng(25).uniform(size=shape).astype(np.float32)
    out = shift_tex(_t(img), *s).cpu().numpy()
    assert _rel(out, oracle.shift_tex(img, *s)) < 1e-6
    outz = shift_z_tex(_t(img), s[2]).cpu().numpy()
    assert _rel(outz, oracle.shift_z_tex(img, s[2])) < 1e-6


@pytest.mark.parametrize("phi", [0.0, 0.3, math.pi / 2, 2.5, -1.1])
def test_rotate_vs_oracle(oracle, phi):
    from nefct_hip.deform import rotate_tex
    shape = (5, 40, 37)
    img = np.random.default_rng(26).uniform(size=shape).astype(np.float32)
    c, s = math.cos(phi), math.sin(phi)
    out = rotate_tex(_t(img), c, s).cpu().numpy()
    assert _rel(out, oracle.rotate_tex(img, c, s)) < 1e-6


def test_reference_drop_in_ops(oracle):
    """The deform_tex op names and argument order, volumes (nx, ny, nz) x fastest."""
    from nefct_hip import reference as R
    shape = (9, 12, 14)   # nz, ny, nx
    nz, ny, nx = shape
    grid = [nx, ny, nz]
    img = np.random.default_rng(27).uniform(size=shape).astype(np.float32)
    m = _smooth_field(shape, 28, amp=1.5)
    as_xyz = lambda a: _t(a).reshape(nx, ny, nz)   # same memory, reference's (x, y, z) shape
    out = R.deform_tex.Deform(as_xyz(img), *(as_xyz(a) for a in m), grid)
    assert tuple(out.shape) == (nx, ny, nz)
    assert _rel(out.reshape(shape).cpu().numpy(), oracle.deform_tex(img, *m)) < 1e-6
    out = R.deform_tex.DeformInvert(as_xyz(img), *(as_xyz(a) for a in m), grid)
    assert _rel(out.reshape(shape).cpu().numpy(), oracle.deform_invert_tex(img, *m)) < 1e-5
    out = R.deform_tex.Shift(as_xyz(img), grid, 0.5, -1.5, 2.0)
    assert _rel(out.reshape(shape).cpu().numpy(), oracle.shift_tex(img, 0.5, -1.5, 2.0)) < 1e-6
    out = R.deform_tex.ShiftZ(as_xyz(img), grid, -1.25)
    assert _rel(out.reshape(shape).cpu().numpy(), oracle.shift_z_tex(img, -1.25)) < 1e-6
    out = R.deform_tex.Rotate(as_xyz(img), grid, math.cos(0.4), math.sin(0.4))
    assert _rel(out.reshape(shape).cpu().numpy(), oracle.rotate_tex(img, math.cos(0.4), math.sin(0.4))) < 1e-6


def test_nonfinite_displacement_and_errors():
    from nefct_hip import NefctError
    from nefct_hip.deform import deform_tex, invert_dvf
    shape = (6, 7, 8)
    img = torch.rand(shape, device=DEV)
    z = torch.zeros(shape, device=DEV)
    mx = z.clone()
    mx[2, 3, 4] = float("nan")
    out = deform_tex(img, mx, z, z)
    assert torch.isnan(out[2, 3, 4])
    mask = torch.ones(shape, dtype=torch.bool, device=DEV)
    mask[2, 3, 4] = False
    assert torch.equal(out[mask], img[mask])
    inf = z.clone()
    inf[0, 0, 0] = float("inf")
    out = deform_tex(img, inf, z, z)
    assert out[0, 0, 0].item() == img[0, 0, -1].item()   # +inf clamps to the last x texel
    with pytest.raises(NefctError):
        deform_tex(img, z[:-1], z, z)
    with pytest.raises(NefctError):
        invert_dvf(z, z, z, -1)
    empty = torch.zeros((0, 4, 4), device=DEV)
    with pytest.raises(NefctError):
        deform_tex(empty, empty, empty, empty)
