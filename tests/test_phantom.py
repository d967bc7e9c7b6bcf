"""Known answers for the ellipsoid rasteriser (nefct/data/sl_phantom.py:109-148):
the voxel volume of an ellipsoid against 4/3 pi a b c, its centroid against
the requested world centre, and the direction of its long axis under a
general ZXZ Euler triple against column 0 of the reference's alpha matrix,
written out here from the rotation's definition: rotate by psi about z, then
theta about x, then phi about z, alpha = R_z(psi) R_x(theta) R_z(phi) in the
passive (frame) form, and the frame map x' = alpha^T g of sl_phantom.py:145."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")


def _rz(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, s, 0.0], [-s, c, 0.0], [0.0, 0.0, 1.0]])


def _rx(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def _moments(img, vol):
    nz, ny, nx = img.shape
    vx, vy, vz = vol.voxel
    z, y, x = np.meshgrid(*[vol.center[i] - vol.size[i] / 2 + (np.arange(n) + 0.5) * v
                            for i, n, v in ((2, nz, vz), (1, ny, vy), (0, nx, vx))], indexing="ij")
    m = img > 0
    pts = np.stack([x[m], y[m], z[m]], axis=1)
    c = pts.mean(axis=0)
    cov = np.cov((pts - c).T)
    w, v = np.linalg.eigh(cov)
    return m.sum() * vx * vy * vz, c, v[:, np.argmax(w)]


def test_euler_matrix_is_the_passive_zxz_product():
    from nefct_hip.phantom import _zxz
    phi, theta, psi = 0.5, 0.9, -1.2
    want = _rz(psi) @ _rx(theta) @ _rz(phi)
    assert np.allclose(_zxz(phi, theta, psi), want, atol=1e-14)


@pytest.mark.parametrize("euler_deg", [(30.0, 0.0, 0.0), (30.0, 50.0, 20.0), (-70.0, 120.0, 45.0)])
def test_ellipsoid_volume_centre_and_long_axis(euler_deg):
    from nefct_hip import VolumeGeometry
    from nefct_hip.phantom import rasterise_ellipsoids
    vol = VolumeGeometry.cube((112, 104, 96), 1.0, center=(3.0, -2.0, 1.0))
    axes = (40.0, 14.0, 9.0)
    centre = (6.5, -4.25, 3.0)
    euler = tuple(math.radians(a) for a in euler_deg)
    img = rasterise_ellipsoids(vol, [(axes, 0.02, centre, euler)], dtype=torch.float64).numpy()
    assert set(np.unique(img)) == {0.0, 0.02}
    volume, c, axis = _moments(img, vol)
    assert volume == pytest.approx(4.0 / 3.0 * math.pi * axes[0] * axes[1] * axes[2], rel=0.01)
    assert np.allclose(c, np.add(centre, vol.center), atol=0.05)
    alpha = _rz(euler[2]) @ _rx(euler[1]) @ _rz(euler[0])
    want = alpha[:, 0]          # x' = alpha^T g: semi-axis 0 along column 0
    assert abs(float(np.dot(axis, want))) > math.cos(math.radians(0.5))
    # the transposed convention (semi-axis 0 along row 0) points elsewhere
    if euler_deg != (0.0, 0.0, 0.0):
        assert abs(float(np.dot(axis, alpha[0, :]))) < 0.99


def test_seeded_phantom_is_reproducible_and_additive():
    from nefct_hip import VolumeGeometry
    from nefct_hip.phantom import ellipsoid_params, ellipsoid_phantom, rasterise_ellipsoids
    vol = VolumeGeometry.cube((48, 48, 40), 2.0)
    a = ellipsoid_phantom(vol, 7, 5, dtype=torch.float64)
    b = ellipsoid_phantom(vol, 7, 5, dtype=torch.float64)
    assert torch.equal(a, b)
    parts = sum(rasterise_ellipsoids(vol, [p], dtype=torch.float64) for p in ellipsoid_params(7, 5, vol.size))
    assert torch.allclose(a, parts, atol=1e-15)
