"""Pins of the oracle's restatements of the voxel-driven backprojector
(nefct/cc/pixel_3d/pixel.cu.cc:12-124) and of the pixel-driven footprint FDK
backprojector (nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255).

The reference ships no tests or vectors for either, so the pins are:
* voxel-driven: a sinogram that is linear in the detector indices is
  reproduced exactly by bilinear sampling, so the backprojection of a ramp is
  the sum over views of the ramp at the voxel's projection, and that
  projection is computed here independently, by intersecting the line from
  the source through the voxel centre with the detector of the Siddon op's
  ray geometry (siddon.cu.cc:341-353, via the oracle's ray_points);
* footprint FDK: per pixel and slab the overlap weights partition the
  footprint (sum 1 once the distance weight is divided out), and at angle 0
  the footprint is exactly the pixel's shadow, so its centroid is on the
  pixel's centre ray.
"""
import math

import numpy as np
import pytest


def _siddon_pixel_coords(oracle, det, angle, P):
    """Fractional Siddon-pixel coordinates (u, w) of the projection of point P
    onto the detector of the Siddon op geometry; pixel centres at integers."""
    S, D00 = oracle.ray_points(det, angle, 0, 0)
    _, D10 = oracle.ray_points(det, angle, 1, 0)
    _, D01 = oracle.ray_points(det, angle, 0, 1)
    S, D00, D10, D01 = (np.asarray(v, np.float64) for v in (S, D00, D10, D01))
    ea, eb = D10 - D00, D01 - D00
    # S + t (P - S) = D00 + u ea + w eb
    A = np.stack([P - S, -ea, -eb], axis=1)
    t, u, w = np.linalg.solve(A, D00 - S)
    return u, w


@pytest.mark.parametrize("na,nb", [(40, 30), (41, 31)])
def test_pixel_bp_linear_ramp_matches_ray_geometry(oracle, na, nb):
    """Flat detector.  pv = alpha*ia + beta*ib + gamma*iv is reproduced by the
    bilinear sample, so BP(voxel) = sum_v alpha a_v + beta b_v + gamma v, with
    (a_v, b_v) = Siddon-pixel coordinates shifted by the pixel op's index
    convention: a = u - 0.5 + (na//2 - na/2) + (ai_siddon - ai) (Q7)."""
    SID, SAD, da, db, ai, bi = 300.0, 480.0, 2.0, 2.2, 0.3, -0.4
    det = oracle.det_struct(0, SID, SAD, na, da, ai, nb, db, bi)
    grid = (6, 5, 4)
    size = (12.0, 10.0, 8.0)
    low = (-6.0, -5.0, -4.0)
    angles = np.array([0.0, 0.7, 2.1, 4.0])
    alpha, beta, gamma = 0.37, -0.21, 0.05
    iv, ib, ia = np.meshgrid(np.arange(4), np.arange(nb), np.arange(na), indexing="ij")
    pv = alpha * ia + beta * ib + gamma * iv
    out = oracle.pixel_bp(0, grid, low, size, angles, SID, SAD, na, da, ai, nb, db, bi, pv)
    for iz in range(grid[2]):
        for iy in range(grid[1]):
            for ix in range(grid[0]):
                P = np.array([low[0] + (ix + 0.5) * size[0] / grid[0], low[1] + (iy + 0.5) * size[1] / grid[1],
                              low[2] + (iz + 0.5) * size[2] / grid[2]])
                want = 0.0
                for v, th in enumerate(angles):
                    u, w = _siddon_pixel_coords(oracle, det, th, P)
                    a = u - 0.5 + (na // 2 - na / 2)
                    b = w - 0.5 + (nb // 2 - nb / 2)
                    assert 0 <= a < na - 1 and 0 <= b < nb - 1
                    want += alpha * a + beta * b + gamma * v
                assert out[iz, iy, ix] == pytest.approx(want, rel=1e-10, abs=1e-9)


def test_pixel_bp_cylindrical_angle_and_flat_z(oracle):
    """Cylindrical detector: the in-plane index is the voxel's angle about the
    source (the Siddon op's cylinder pixels are equal angles about the source,
    siddon.cu.cc:376-387), while z is magnified as on a flat detector at
    distance SAD (pixel.cu.cc:72, not the cylinder's own height).  Both are
    read off the Siddon geometry of pixels (0, 0), (1, 0) and (0, 1)."""
    SID, SAD, na, da, nb, db = 300.0, 480.0, 50, 0.01, 30, 2.0
    det = oracle.det_struct(1, SID, SAD, na, da, 0.0, nb, db, 0.0)
    grid, size, low = (3, 3, 3), (6.0, 6.0, 6.0), (-3.0, -3.0, -3.0)
    th = 0.4
    S, D00 = (np.asarray(v) for v in oracle.ray_points(det, th, 0, 0))
    _, D10 = (np.asarray(v) for v in oracle.ray_points(det, th, 1, 0))
    _, D01 = (np.asarray(v) for v in oracle.ray_points(det, th, 0, 1))
    centre = -S / np.linalg.norm(S)                     # direction source -> axis

    def ang_about_source(v):
        d = v - S
        return math.atan2(centre[0] * d[1] - centre[1] * d[0], centre[0] * d[0] + centre[1] * d[1])

    a0, a1 = ang_about_source(D00), ang_about_source(D10)
    iv, ib, ia = np.meshgrid(np.arange(1), np.arange(nb), np.arange(na), indexing="ij")
    for alpha, beta in ((1.0, 0.0), (0.0, 1.0)):
        pv = alpha * ia + beta * ib + 0.0
        out = oracle.pixel_bp(1, grid, low, size, [th], SID, SAD, na, da, 0.0, nb, db, 0.0, pv)
        for (iz, iy, ix) in [(0, 0, 0), (2, 1, 2), (1, 2, 0)]:
            P = np.array([low[0] + (ix + 0.5) * 2.0, low[1] + (iy + 0.5) * 2.0, low[2] + (iz + 0.5) * 2.0])
            u = (ang_about_source(P) - a0) / (a1 - a0)      # Siddon pixel index, centres at integers
            d = P - S
            dist = centre[0] * d[0] + centre[1] * d[1]      # depth along the central ray
            z_flat = P[2] * SAD / dist                      # height on a flat detector at SAD
            w = (z_flat - D00[2]) / (D01[2] - D00[2])
            want = alpha * (u - 0.5) + beta * (w - 0.5)
            assert out[iz, iy, ix] == pytest.approx(want, rel=1e-9, abs=1e-9)


def test_pixel_bp_skips_views_off_the_detector(oracle):
    SID, SAD, na, da, nb, db = 100.0, 150.0, 4, 1.0, 4, 1.0
    pv = np.ones((1, nb, na))
    out = oracle.pixel_bp(0, (3, 1, 1), (-30.0, 20.0, -0.5), (60.0, 1.0, 1.0), [0.0], SID, SAD, na, da, 0.0, nb,
                          db, 0.0, pv)
    # y = 20.5 mm projects far off a 4-pixel detector
    assert np.all(out == 0.0)
    out = oracle.pixel_bp(0, (1, 1, 1), (-0.5, -0.5, -0.5), (1.0, 1.0, 1.0), [0.0], SID, SAD, na, da, 0.0, nb, db,
                          0.0, pv)
    assert out[0, 0, 0] == pytest.approx(1.0)


def _fp_geom():
    # voxel units, dist_3d_back_fdk.cu.cc naming: SAD = source-axis, SID = source-detector
    return dict(SID=90.0, SAD=60.0, na=24, da=1.3, ai=0.0, nb=10, db=1.1, bi=0.0)


@pytest.mark.parametrize("angle", [0.0, 0.35, math.pi / 2, 2.5])
def test_footprint_weights_partition_each_slab(oracle, angle):
    """One lit pixel: dividing each voxel's value by its distance weight
    |r|/|r_dom| SAD^2/U^2 leaves wt * wz, which sums to 1 over every slab whose
    footprint lies inside the volume."""
    g = _fp_geom()
    nx = ny = 40
    nz = 24
    na, nb = g["na"], g["nb"]
    for ia, ib in [(11, 4), (3, 7), (20, 1)]:
        pv = np.zeros((1, nb, na))
        pv[0, ib, ia] = 1.0
        out = oracle.fdk_fp_bp(0, (nx, ny, nz), [angle], None, g["SID"], g["SAD"], na, g["da"], g["ai"], nb,
                               g["db"], g["bi"], pv)
        c, s = math.cos(angle), math.sin(angle)
        x1, y1 = -g["SAD"] * c, -g["SAD"] * s
        x20 = g["SID"] - g["SAD"]
        y20 = (ia + 0.5) * g["da"] - na * g["da"] / 2
        z21 = (ib + 0.5) * g["db"] - nb * g["db"] / 2
        x21, y21 = x20 * c - y20 * s - x1, x20 * s + y20 * c - y1
        xdom = abs(x21) > abs(y21)
        rdom = x21 if xdom else y21
        ln = math.sqrt(x21 ** 2 + y21 ** 2 + z21 ** 2)
        iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        u = g["SAD"] + (ix + 0.5 - nx // 2) * c + (iy + 0.5 - ny // 2) * s
        w = out / (ln / abs(rdom) * g["SAD"] ** 2 / u ** 2)
        per_slab = w.sum(axis=(0, 1)) if xdom else w.sum(axis=(0, 2))
        # slabs where the centre ray is at least 3 cells inside the volume (the
        # footprint is about 2 cells wide here)
        n_m, n_t = (nx, ny) if xdom else (ny, nx)
        src_m, src_t = (x1, y1) if xdom else (y1, x1)
        rt = y21 if xdom else x21
        mc = np.arange(n_m) + 0.5 - n_m / 2 - src_m
        t_c = rt / rdom * mc + src_t + n_t // 2
        z_c = z21 / rdom * mc + nz // 2
        inside = (t_c >= 3) & (t_c <= n_t - 3) & (z_c >= 3) & (z_c <= nz - 3)
        assert inside.sum() >= 10
        assert np.allclose(per_slab[inside], 1.0, atol=1e-12), (ia, ib, per_slab[inside])
        # outside the volume nothing is deposited
        assert np.all(per_slab <= 1.0 + 1e-12)


@pytest.mark.parametrize("cyl", [0, 1])
def test_footprint_is_the_pixel_shadow_at_angle_zero(oracle, cyl):
    """At angle 0 the pixel's shadow on a slab mid-plane at distance m from
    the source is bounded by its edge rays: [min, max] of (y_edge / x_edge) m
    in y and (z_centre +- db/2) / x_centre m in z, shifted by n // 2.  Every
    cell then carries (its overlap with that interval) / (interval length),
    and z and y factorise."""
    g = _fp_geom()
    if cyl:
        g["da"] = 0.02
    nx = ny = 40
    nz = 24
    na, nb = g["na"], g["nb"]
    ia, ib = 9, 6
    pv = np.zeros((1, nb, na))
    pv[0, ib, ia] = 1.0
    out = oracle.fdk_fp_bp(cyl, (nx, ny, nz), [0.0], None, g["SID"], g["SAD"], na, g["da"], g["ai"], nb, g["db"],
                           g["bi"], pv)

    def edge(e):
        # detector point of pixel edge e relative to the source (source at -SAD on x)
        if cyl:
            ang = e * g["da"] - na * g["da"] / 2
            return g["SID"] * math.cos(ang), g["SID"] * math.sin(ang)
        return g["SID"], e * g["da"] - na * g["da"] / 2

    (xl, yl), (xr, yr), (xc_, yc_) = edge(ia), edge(ia + 1), edge(ia + 0.5)
    z21 = (ib + 0.5) * g["db"] - nb * g["db"] / 2
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = g["SAD"] + (ix + 0.5 - nx // 2)
    ln = math.sqrt(xc_ ** 2 + yc_ ** 2 + z21 ** 2)
    w = out / (ln / abs(xc_) * g["SAD"] ** 2 / u ** 2)

    def overlaps(lo, hi, n):
        cells = np.arange(n)
        return np.clip(np.minimum(cells + 1.0, hi) - np.maximum(cells + 0.0, lo), 0.0, None) / (hi - lo)

    for ixs in range(10, 30):
        m = ixs + 0.5 - nx / 2 + g["SAD"]      # distance of the slab mid-plane from the source
        ys = [yl / xl * m, yl / xr * m, yr / xl * m, yr / xr * m]
        t1, t2 = min(ys) + ny // 2, max(ys) + ny // 2
        z1, z2 = sorted([(z21 - g["db"] / 2) / xc_ * m + nz // 2, (z21 + g["db"] / 2) / xc_ * m + nz // 2])
        want = overlaps(z1, z2, nz)[:, None] * overlaps(t1, t2, ny)[None, :]
        assert np.allclose(w[:, :, ixs], want, rtol=1e-10, atol=1e-13)
