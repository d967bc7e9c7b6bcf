"""Pins the oracle's warp, FDK and TV restatements (numpy, FP64) against
mathematics that fixes them, and the product's filter-tap design against the
restated MATLAB filter.

* deform (cpp/deform/deform.cu.cc:13-46): identity and integer shifts, exact
  reproduction of affine fields on linear images, transpose by dot test,
  displacement gradient by central differences.
* FDK weights/filter (mat_codes/filtering.m, ParkerWeight.m): the ramp has no
  DC response; Parker weights of conjugate rays sum to one; the spatial taps
  used by the GPU filter equal the zero-padded FFT filter; reconstructing a
  uniform ball from exact projections returns its attenuation at the centre.
* TV (nefct/functions/total_variation.py): adjoint by dot test.
"""
import math

import numpy as np
import pytest


def test_deform_identity_and_shift(oracle):
    rng = np.random.default_rng(0)
    img = rng.uniform(size=(7, 8, 9))
    z = np.zeros_like(img)
    out = oracle.deform(img, z, z, z, boundary="ref")
    inner = (slice(1, -1),) * 3
    assert np.allclose(out[inner], img[inner], atol=1e-12)
    # reference rule: base index 0 and n-1 are never written
    assert np.all(out[0] == 0) and np.all(out[:, :, -1] == 0)
    out0 = oracle.deform(img, z, z, z, boundary="zero")
    assert np.allclose(out0, img, atol=1e-12)
    # integer shift by +1 in x: out[.., ix] = img[.., ix + 1]
    one = np.ones_like(img)
    out = oracle.deform(img, one, z, z, boundary="zero")
    assert np.allclose(out[:, :, :-1], img[:, :, 1:], atol=1e-12)
    assert np.all(out[:, :, -1] == 0)


def test_deform_affine_field_exact(oracle):
    nz, ny, nx = 10, 11, 12
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    img = 0.5 * ix - 0.25 * iy + 0.75 * iz + 2.0
    mx = 0.3 + 0.01 * iy
    my = -0.2 + 0.02 * iz
    mz = 0.15 - 0.01 * ix
    out = oracle.deform(img, mx, my, mz, boundary="ref")
    expect = 0.5 * (ix + mx) - 0.25 * (iy + my) + 0.75 * (iz + mz) + 2.0
    # in float32 the sample position carries ~1e-6 rounding
    mask = out != 0
    assert mask.sum() > 0.5 * out.size
    assert np.allclose(out[mask], expect[mask], atol=2e-5)


@pytest.mark.parametrize("boundary", ["ref", "zero"])
def test_deform_adjoint_dot(oracle, boundary):
    rng = np.random.default_rng(1)
    shape = (9, 10, 11)
    mx, my, mz = (rng.uniform(-2, 2, size=shape) for _ in range(3))
    x = rng.normal(size=shape)
    y = rng.normal(size=shape)
    lhs = np.vdot(oracle.deform(x, mx, my, mz, boundary), y)
    rhs = np.vdot(x, oracle.deform_adjoint(y, mx, my, mz, boundary))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("boundary", ["ref", "zero"])
def test_deform_dvf_grad_finite_differences(oracle, boundary):
    rng = np.random.default_rng(2)
    shape = (8, 9, 10)
    img = rng.uniform(size=shape)
    g = rng.normal(size=shape)
    # fractional parts in [0.2, 0.8] keep +-eps inside one trilinear cell
    mx, my, mz = (np.floor(rng.uniform(-2, 2, size=shape)) + rng.uniform(0.2, 0.8, size=shape) for _ in range(3))
    gx, gy, gz = oracle.deform_dvf_grad(img, g, mx, my, mz, boundary)
    # the interpolant is linear in each coordinate inside a cell, so a central
    # difference that stays in the cell is exact up to rounding
    eps = 0.1
    for comp, grad in ((0, gx), (1, gy), (2, gz)):
        m = [mx, my, mz]
        p = [a.copy() for a in m]
        q = [a.copy() for a in m]
        p[comp] = p[comp] + eps
        q[comp] = q[comp] - eps
        fd = (oracle.deform(img, *p, boundary) - oracle.deform(img, *q, boundary)) * g / (2 * eps)
        assert np.allclose(fd, grad, rtol=1e-5, atol=1e-6), comp


def test_ramp_has_no_dc_and_matches_textbook_samples():
    from oracle import oracle as O
    r = O.ramp_response(512)
    assert r[0] < 1e-3 * r.max()
    # Ram-Lak: response grows linearly with frequency (|w|/pi) away from DC
    n = r.size
    w = 2 * np.pi * np.arange(n // 2 + 1) / n
    assert np.allclose(r[n // 8: n // 2], (w / np.pi)[n // 8: n // 2], rtol=0.05)


def test_parker_conjugate_rays_sum_to_one(oracle):
    """A ray at p = beta + gamma and its conjugate at p + 180 (same line, the
    other way round) carry weights summing to one over the whole short scan."""
    fan = 23.1
    # open interval: at exactly p = fan/2 the MATLAB region tests give 1 + 1
    # (ParkerWeight.m:11 uses <= while :16 uses <), a measure-zero seam
    p = np.linspace(-fan / 2, fan / 2, 20001)[1:-1]
    w = oracle.parker_w(p, fan) + oracle.parker_w(p + 180.0, fan)
    assert np.allclose(w, 1.0, atol=1e-12)
    q = np.linspace(fan / 2, 180 - fan / 2, 5001)[1:-1]   # middle: weight 1, conjugate 0
    assert np.allclose(oracle.parker_w(q, fan) + oracle.parker_w(q + 180.0, fan), 1.0)
    # the table is that function of beta + gamma
    det = oracle.det_struct(0, 1000.0, 1500.0, 64, 6.0, 0.5, 8, 6.0, 0.5)
    u = oracle.detector_coords(64, 6.0, 0.5)
    gam = np.degrees(np.arctan(u / 1500.0))
    F = abs(gam[-1] - gam[0]) / 2
    nv, dang = 400, 0.5
    tab = oracle.parker_weights(nv, np.radians(dang), det)
    beta = ((180 / dang - nv) / 2.0 + np.arange(1, nv + 1)) * dang
    assert np.allclose(tab, oracle.parker_w(beta[:, None] + gam[None, :], F), atol=1e-12)
    # smooth transitions: weights stay in [0, 1]
    assert tab.min() >= 0 and tab.max() <= 1


def test_product_filter_taps_equal_zero_padded_fft_filter(oracle):
    from nefct_hip.fdk import ramp_taps
    rng = np.random.default_rng(3)
    for na, win in ((100, "ram-lak"), (257, "hann"), (64, "shepp-logan"), (90, "hamming"), (33, "cosine")):
        rows = rng.normal(size=(3, na))
        want = oracle.ramp_filter_rows(rows, win)
        taps = ramp_taps(na, win).astype(np.float64)
        got = np.stack([np.convolve(r, taps)[na - 1: 2 * na - 1] for r in rows])
        assert np.allclose(got, want, atol=1e-5 * np.abs(want).max()), (na, win)


def _fdk_oracle(oracle, vol, det, angles, sino, parker=False):
    nv = len(angles)
    dang = abs(angles[1] - angles[0])
    q = sino * oracle.cosine_weight(det)[None]
    if parker:
        q = q * oracle.parker_weights(nv, dang, det)[:, None, :]
    q = oracle.ramp_filter_rows(q) * oracle.fdk_scale(det, nv, dang, parker)
    return oracle.fdk_backproject(vol, det, angles, q)


def test_fdk_uniform_ball_full_scan(oracle):
    n = 32
    V = oracle.vol_struct((n, n, n), (0, 0, 0), (n * 2.0,) * 3)
    g = (np.arange(n) + 0.5) * 2.0 - n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    mu = 0.02
    img = mu * ((X ** 2 + Y ** 2 + Z ** 2) <= 20.0 ** 2)
    det = oracle.det_struct(0, 300.0, 450.0, 96, 1.5, 0.5, 96, 1.5, 0.5)
    ang = 2 * np.pi * np.arange(120) / 120
    sino = oracle.exact_fwd(V, det, ang, img)
    rec = _fdk_oracle(oracle, V, det, ang, sino)
    centre = rec[n // 2 - 3: n // 2 + 3, n // 2 - 3: n // 2 + 3, n // 2 - 3: n // 2 + 3].mean()
    assert centre == pytest.approx(mu, rel=0.03)
    outside = rec[:, :, :3].mean()
    assert abs(outside) < 0.1 * mu


def test_fdk_uniform_ball_parker_short_scan(oracle):
    n = 32
    V = oracle.vol_struct((n, n, n), (0, 0, 0), (n * 2.0,) * 3)
    g = (np.arange(n) + 0.5) * 2.0 - n
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    mu = 0.02
    img = mu * ((X ** 2 + Y ** 2 + Z ** 2) <= 20.0 ** 2)
    det = oracle.det_struct(0, 300.0, 450.0, 96, 1.5, 0.5, 96, 1.5, 0.5)
    fan = 2 * np.degrees(np.arctan(96 * 1.5 / 2 / 450.0))
    dang = 1.0
    nv = int(np.ceil((180 + fan) / dang)) + 1
    ang = np.radians(dang) * np.arange(nv)
    sino = oracle.exact_fwd(V, det, ang, img)
    rec = _fdk_oracle(oracle, V, det, ang, sino, parker=True)
    centre = rec[n // 2 - 3: n // 2 + 3, n // 2 - 3: n // 2 + 3, n // 2 - 3: n // 2 + 3].mean()
    assert centre == pytest.approx(mu, rel=0.04)


def test_tv_adjoint_dot(oracle):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(6, 7, 8))
    d = rng.normal(size=(3, 6, 7, 8))
    assert np.vdot(oracle.tv_forward(x), d) == pytest.approx(np.vdot(x, oracle.tv_adjoint(d)), rel=1e-12)
