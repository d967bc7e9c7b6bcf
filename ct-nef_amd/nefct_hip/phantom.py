"""Seeded synthetic inputs: ellipsoid phantoms and smooth displacement fields.

The ellipsoid rasteriser follows the 3-D branch of the reference's Shepp-Logan
generator (nefct/data/sl_phantom.py:109-148): indicator of a rotated
ellipsoid (ZXZ Euler angles and the alpha^T frame map of :129-147) sampled at
voxel centres in mm, intensities added.  Parameters are drawn from ``numpy.random.Generator(PCG64(seed))``:
semi-axes U[5, 50] mm, attenuation U[0.01, 0.03] /mm, centres U(+-0.35 size)
per axis, Euler angles U[0, 2 pi).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .geometry import VolumeGeometry


def ellipsoid_params(seed: int, count: int, size_mm):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        axes = rng.uniform(5.0, 50.0, 3)
        mu = rng.uniform(0.01, 0.03)
        centre = np.array([rng.uniform(-0.35 * s, 0.35 * s) for s in size_mm])
        euler = rng.uniform(0.0, 2 * math.pi, 3)
        out.append((axes, mu, centre, euler))
    return out


def _zxz(phi, theta, psi):
    """The ZXZ Euler matrix of sl_phantom.py:137-143 (called alpha there)."""
    cphi, sphi = math.cos(phi), math.sin(phi)
    cth, sth = math.cos(theta), math.sin(theta)
    cpsi, spsi = math.cos(psi), math.sin(psi)
    return np.array([[cpsi * cphi - cth * sphi * spsi, cpsi * sphi + cth * cphi * spsi, spsi * sth],
                     [-spsi * cphi - cth * sphi * cpsi, -spsi * sphi + cth * cphi * cpsi, cpsi * sth],
                     [sth * sphi, -sth * cphi, cth]])


def rasterise_ellipsoids(vol: VolumeGeometry, ellipsoids, device="cpu", dtype=torch.float32):
    """Volume [nz, ny, nx] of additive ellipsoids sampled at voxel centres.

    ``ellipsoids``: iterable of (semi_axes_mm[3], mu, centre_mm[3], euler[3]),
    centres relative to the volume centre.  A world point g (relative to the
    ellipsoid's centre) has ellipsoid-frame coordinates alpha^T g, as in
    sl_phantom.py:145-147 (x' = sum_i g_i alpha[i, 0], ...): semi-axis 0 lies
    along column 0 of alpha, and so on.  Unlike the reference, whose centres
    are given in the rotated frame, centres here are world positions."""
    nx, ny, nz = vol.grid
    vx, vy, vz = vol.voxel
    img = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    x = torch.tensor(vol.center[0] - vol.size[0] / 2 + (np.arange(nx) + 0.5) * vx, dtype=torch.float64, device=device)
    y = torch.tensor(vol.center[1] - vol.size[1] / 2 + (np.arange(ny) + 0.5) * vy, dtype=torch.float64, device=device)
    z = torch.tensor(vol.center[2] - vol.size[2] / 2 + (np.arange(nz) + 0.5) * vz, dtype=torch.float64, device=device)
    for axes, mu, centre, euler in ellipsoids:
        axes = np.asarray(axes, np.float64)
        R = _zxz(*euler)
        r = float(axes.max())
        # bounding box in index space
        lo = [int(max(0, math.floor((centre[i] + vol.center[i] - r - (vol.center[i] - vol.size[i] / 2)) / vs)))
              for i, vs in enumerate((vx, vy, vz))]
        hi = [int(min(n, math.ceil((centre[i] + vol.center[i] + r - (vol.center[i] - vol.size[i] / 2)) / vs) + 1))
              for i, (n, vs) in enumerate(zip((nx, ny, nz), (vx, vy, vz)))]
        if any(h <= l for l, h in zip(lo, hi)):
            continue
        X = x[lo[0]:hi[0]] - (centre[0] + vol.center[0])
        Y = y[lo[1]:hi[1]] - (centre[1] + vol.center[1])
        Z = z[lo[2]:hi[2]] - (centre[2] + vol.center[2])
        Zg, Yg, Xg = torch.meshgrid(Z, Y, X, indexing="ij")
        u = Xg * R[0, 0] + Yg * R[1, 0] + Zg * R[2, 0]
        v = Xg * R[0, 1] + Yg * R[1, 1] + Zg * R[2, 1]
        w = Xg * R[0, 2] + Yg * R[1, 2] + Zg * R[2, 2]
        inside = (u / axes[0]) ** 2 + (v / axes[1]) ** 2 + (w / axes[2]) ** 2 <= 1.0
        img[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += inside.to(torch.float64) * mu
    return img.to(dtype)


def ellipsoid_phantom(vol: VolumeGeometry, seed: int, count: int, device="cpu", dtype=torch.float32):
    """Volume [nz, ny, nx] of ``count`` seeded additive ellipsoids (attenuation in 1/mm)."""
    return rasterise_ellipsoids(vol, ellipsoid_params(seed, count, vol.size), device=device, dtype=dtype)


def gaussian_dvf(shape_zyx, seed: int, voxel_mm=1.0, bumps=8, device="cpu"):
    """Smooth displacement field [3, nz, ny, nx] in voxel units (z component
    only; the warp takes voxel units): sum of Gaussian bumps, sigma U[20, 40]
    voxels, peak U[0, 10] mm, centres uniform.  ``voxel_mm`` is the z voxel
    size (a number, or the (x, y, z) voxel sizes) that turns mm into voxels."""
    vz = float(voxel_mm[2] if isinstance(voxel_mm, (tuple, list)) else voxel_mm)
    if not vz > 0:
        raise ValueError("voxel_mm must be positive")
    rng = np.random.Generator(np.random.PCG64(seed))
    nz, ny, nx = shape_zyx
    iz = torch.arange(nz, dtype=torch.float64, device=device)[:, None, None]
    iy = torch.arange(ny, dtype=torch.float64, device=device)[None, :, None]
    ix = torch.arange(nx, dtype=torch.float64, device=device)[None, None, :]
    dz = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)
    for _ in range(bumps):
        sigma = rng.uniform(20.0, 40.0)
        peak = rng.uniform(0.0, 10.0)
        c = (rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz))
        dz += peak * torch.exp(-((ix - c[0]) ** 2 + (iy - c[1]) ** 2 + (iz - c[2]) ** 2) / (2 * sigma * sigma))
    dvf = torch.zeros((3, nz, ny, nx), dtype=torch.float32, device=device)
    dvf[2] = (dz / vz).to(torch.float32)
    return dvf
