"""Pins of the oracle's texture-rule warps (oracle.tex_sample and the
deform_tex restatements of nefct/cc/deform/deform_tex.cu.cc) by independent
checks: a brute-force per-point sampler written from the texture rule,
texel-centre and clamp identities, integer shifts and quarter-turn rotations
as index permutations, and the closed-form inverse of constant and affine
displacement fields."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _brute_sample(vol, u, v, w):
    """One point at a time: texel i covers [i, i+1) on each axis, value is the
    product-weighted sum over the two nearest texel centres, indices clamped."""
    nz, ny, nx = vol.shape

    def axis(c, n):
        x = c - 0.5
        i = math.floor(x)
        a = x - i
        return [(min(max(i, 0), n - 1), 1.0 - a), (min(max(i + 1, 0), n - 1), a)]

    total = 0.0
    for zi, wz in axis(w, nz):
        for yi, wy in axis(v, ny):
            for xi, wx in axis(u, nx):
                total += vol[zi, yi, xi] * wx * wy * wz
    return total


def test_tex_sample_matches_brute_force():
    rng = np.random.default_rng(11)
    vol = rng.standard_normal((4, 5, 6))
    pts = rng.uniform(-3.0, 9.0, size=(400, 3))
    got = O.tex_sample(vol, pts[:, 0], pts[:, 1], pts[:, 2])
    want = np.array([_brute_sample(vol, *p) for p in pts])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_tex_sample_centres_midpoints_and_clamp():
    rng = np.random.default_rng(12)
    vol = rng.standard_normal((3, 4, 5))
    iz, iy, ix = np.meshgrid(np.arange(3), np.arange(4), np.arange(5), indexing="ij")
    np.testing.assert_array_equal(O.tex_sample(vol, ix + 0.5, iy + 0.5, iz + 0.5), vol)
    # halfway between two x texels: their mean
    mid = O.tex_sample(vol, np.full(3, 2.0), np.full(3, 1.5), np.arange(3) + 0.5)
    np.testing.assert_allclose(mid, 0.5 * (vol[:, 1, 1] + vol[:, 1, 2]), atol=1e-14)
    # far outside: the edge texel (clamp addressing)
    far = O.tex_sample(vol, np.array([-50.0, 80.0]), np.array([-7.0, 99.0]), np.array([-1e3, 1e3]))
    np.testing.assert_array_equal(far, [vol[0, 0, 0], vol[-1, -1, -1]])


def test_deform_tex_identity_and_constant_field_equals_shift():
    rng = np.random.default_rng(13)
    img = rng.standard_normal((6, 7, 8)).astype(np.float32)
    z = np.zeros_like(img)
    np.testing.assert_array_equal(O.deform_tex(img, z, z, z), img)
    a, b, c = 1.25, -2.5, 0.75
    got = O.deform_tex(img, z - a, z - b, z - c)
    np.testing.assert_array_equal(got, O.shift_tex(img, a, b, c))


@pytest.mark.parametrize("s", [(1, 0, 0), (0, -2, 0), (0, 0, 3), (2, 1, -1), (40, -40, 0)])
def test_integer_shift_is_edge_replicating_index_shift(s):
    rng = np.random.default_rng(14)
    img = rng.standard_normal((5, 6, 7))
    nz, ny, nx = img.shape
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    want = img[np.clip(iz - s[2], 0, nz - 1), np.clip(iy - s[1], 0, ny - 1), np.clip(ix - s[0], 0, nx - 1)]
    np.testing.assert_array_equal(O.shift_tex(img, *s), want)
    if s[0] == 0 and s[1] == 0:
        np.testing.assert_array_equal(O.shift_z_tex(img, s[2]), want)


def test_rotation_quarter_turns_are_permutations():
    rng = np.random.default_rng(15)
    img = rng.standard_normal((3, 8, 8))
    np.testing.assert_array_equal(O.rotate_tex(img, 1.0, 0.0), img)
    np.testing.assert_array_equal(O.rotate_tex(img, 0.0, 1.0), np.rot90(img, 1, axes=(1, 2)))
    np.testing.assert_array_equal(O.rotate_tex(img, -1.0, 0.0), img[:, ::-1, ::-1])
    # odd sizes keep the identity (the integer centre cancels)
    odd = rng.standard_normal((2, 7, 9))
    np.testing.assert_array_equal(O.rotate_tex(odd, 1.0, 0.0), odd)


def test_invert_constant_field():
    shape = (5, 6, 7)
    m = [np.full(shape, v) for v in (0.7, -1.3, 2.2)]
    x, y, z = O.invert_dvf(*m, iters=30)
    # (the eight weights of a sample sum to 1 up to rounding)
    np.testing.assert_allclose(x, -m[0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(y, -m[1], rtol=0, atol=1e-14)
    np.testing.assert_allclose(z, -m[2], rtol=0, atol=1e-14)
    rng = np.random.default_rng(16)
    img = rng.standard_normal(shape)
    mi = [np.full(shape, v) for v in (1.0, -2.0, 1.0)]
    np.testing.assert_allclose(O.deform_invert_tex(img, *mi), O.shift_tex(img, 1.0, -2.0, 1.0), rtol=0, atol=1e-13)


def test_invert_affine_field_closed_form():
    # m(p) = A p + b is reproduced exactly by trilinear sampling away from the
    # clamped border; the fixed point solves x = -(A (p + x) + b)
    shape = (16, 18, 20)   # nz, ny, nx
    A = np.array([[0.04, -0.02, 0.01], [0.015, -0.03, 0.02], [-0.01, 0.025, 0.05]])
    b = np.array([0.8, -0.6, 0.4])
    nz, ny, nx = shape
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.stack([ix, iy, iz], -1).astype(np.float64)
    M = P @ A.T + b
    x, y, z = O.invert_dvf(M[..., 0], M[..., 1], M[..., 2], iters=30)
    want = -np.linalg.solve(np.eye(3) + A, (P @ A.T + b).reshape(-1, 3).T).T.reshape(P.shape)
    inner = (slice(3, -3),) * 3
    got = np.stack([x, y, z], -1)
    np.testing.assert_allclose(got[inner], want[inner], rtol=0, atol=1e-12)


@pytest.mark.parametrize("iters", [1, 2, 3])
def test_invert_updates_components_in_order(iters):
    # for an affine field the k-th iterate has a closed form; each component
    # is sampled at the point that already holds the components updated
    # before it in the same iteration (x, then y, then z)
    shape = (12, 12, 12)
    A = np.array([[0.06, -0.04, 0.03], [0.05, -0.02, 0.04], [-0.03, 0.05, 0.07]])
    b = np.array([0.5, -0.4, 0.3])
    iz, iy, ix = np.meshgrid(*(np.arange(12),) * 3, indexing="ij")
    P = np.stack([ix, iy, iz], -1).astype(np.float64)
    M = P @ A.T + b
    d = np.zeros_like(P)
    for _ in range(iters):
        for k in range(3):
            d[..., k] = -((P + d) @ A[k] + b[k])
    x, y, z = O.invert_dvf(M[..., 0], M[..., 1], M[..., 2], iters=iters)
    inner = (slice(3, -3),) * 3
    np.testing.assert_allclose(np.stack([x, y, z], -1)[inner], d[inner], rtol=0, atol=1e-12)
