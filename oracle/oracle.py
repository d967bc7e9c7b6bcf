"""CPU oracle for the cone-beam projector path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, by
``__graft_entry__.smoke()`` (as the checker) and by ``bench.py``'s
``cpu_baseline`` leg.  The product (``ct-nef_amd/nefct_hip``) never imports
anything from here.

Contents
--------
* ctypes bindings to ``oracle.c`` (built as ``oracle/liboracle.so`` by
  ``oracle/Makefile``):
    - ``exact_fwd / exact_bp / exact_count / exact_ray`` -- FP64 exact Siddon
      (merged-alpha), forward, exact transpose and intersection count.
    - ``ref_fwd / ref_bp`` -- FP32 restatement of the reference slab walk,
      ``nefct/cc/siddon_3d/siddon.cu.cc:12-321`` and ``:450-764``, with the
      quirk toggles Q1/Q2/Q3/Q6.
* numpy restatements (FP64):
    - ``deform_ref`` -- ``cpp/deform/deform.cu.cc:13-46`` with the zero-filled
      output of ``nefct/cc/deform/deform.cc:94``; ``deform_zero`` (zero-padded
      boundary); their exact transpose and DVF gradient.
    - FDK pre-weights: cosine weight ``mat_codes/filtering.m:5-14``, Parker
      weights ``mat_codes/ParkerWeight.m:5-27``; ramp filter
      ``mat_codes/filtering.m:20-48,54-88``; voxel-driven weighted
      backprojection ``mat_codes/backprojection.m:11-25`` with the
      ``(SOD/U)^2`` weight also used by
      ``nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:154``.
    - anisotropic TV ``nefct/functions/total_variation.py:19-35`` and its
      adjoint ``:125-149``.
    - texture-rule warps ``nefct/cc/deform/deform_tex.cu.cc``: ``tex_sample``
      (clamp-to-edge trilinear), ``deform_tex``, ``invert_dvf`` (fixed-point
      inverse field), ``deform_invert_tex``, ``shift_tex``, ``shift_z_tex``,
      ``rotate_tex``.

Parity status of each function is stated in its docstring; the checks that pin
it live in ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

INV_L2 = 1      # Q1
REF_SLAB = 2    # Q2
SHORT_SKIP = 4  # Q3
TIE_GT = 8      # Q6
REF_FWD_FLAGS = INV_L2 | REF_SLAB
REF_BP_FLAGS = REF_SLAB | SHORT_SKIP | TIE_GT


class _Vol(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("cz", ctypes.c_double),
                ("sx", ctypes.c_double), ("sy", ctypes.c_double), ("sz", ctypes.c_double)]


class _Det(ctypes.Structure):
    _fields_ = [("cyl", ctypes.c_int), ("sid", ctypes.c_double), ("sad", ctypes.c_double),
                ("na", ctypes.c_int), ("da", ctypes.c_double), ("ai", ctypes.c_double),
                ("nb", ctypes.c_int), ("db", ctypes.c_double), ("bi", ctypes.c_double)]


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d, f, i = ctypes.c_double, ctypes.c_float, ctypes.c_int
        L.orc_exact_fwd.argtypes = [P(_Vol), P(_Det), P(d), i, P(d), P(d)]
        L.orc_exact_bp.argtypes = [P(_Vol), P(_Det), P(d), i, P(d), P(d)]
        L.orc_exact_count.argtypes = [P(_Vol), P(_Det), P(d), i, P(ctypes.c_int32)]
        L.orc_exact_count.restype = ctypes.c_long
        L.orc_exact_ray.argtypes = [P(_Vol), P(d), P(d), P(d)]
        L.orc_exact_ray.restype = d
        L.orc_exact_ray_count.argtypes = [P(_Vol), P(d), P(d), P(d)]
        L.orc_exact_ray_count.restype = ctypes.c_long
        L.orc_ray_points.argtypes = [P(_Det), d, i, i, P(d), P(d)]
        L.orc_ref_fwd.argtypes = [P(_Vol), P(_Det), P(f), i, P(f), P(f), ctypes.c_uint, i, i]
        L.orc_ref_bp.argtypes = [P(_Vol), P(_Det), P(f), i, P(f), P(f), ctypes.c_uint, i, i]
        L.orc_ref_fwd_f64.argtypes = [P(_Vol), P(_Det), P(d), i, P(d), P(d), ctypes.c_uint, i, i]
        L.orc_ref_bp_f64.argtypes = [P(_Vol), P(_Det), P(d), i, P(d), P(d), ctypes.c_uint, i, i]
        L.orc_threads.restype = i
        L.orc_pixel_bp.argtypes = [i, P(i), P(d), P(d), P(d), i, d, d, i, d, d, i, d, d, P(d), P(d)]
        L.orc_fdk_fp_bp.argtypes = [i, i, i, i, P(d), P(d), i, d, d, i, d, d, i, d, d, P(d), P(d)]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def vol_struct(shape_xyz, center, size):
    nx, ny, nz = (int(v) for v in shape_xyz)
    return _Vol(nx, ny, nz, *(float(c) for c in center), *(float(s) for s in size))


def det_struct(cyl, SID, SAD, na, da, ai, nb, db, bi):
    return _Det(int(bool(cyl)), float(SID), float(SAD), int(na), float(da), float(ai),
                int(nb), float(db), float(bi))


def threads() -> int:
    return int(lib().orc_threads())


# --------------------------------------------------------------------------
# FP64 exact Siddon
# --------------------------------------------------------------------------
def exact_fwd(vol, det, angles, img_zyx):
    """FP64 exact line integrals, shape (nv, nb, na).  img is [nz, ny, nx]."""
    img = np.ascontiguousarray(img_zyx, dtype=np.float64)
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros((len(ang), det.nb, det.na), np.float64)
    lib().orc_exact_fwd(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_double), len(ang),
                        _ptr(img, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def exact_bp(vol, det, angles, pv):
    """Exact transpose of exact_fwd; returns [nz, ny, nx] float64."""
    p = np.ascontiguousarray(pv, dtype=np.float64)
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros((vol.nz, vol.ny, vol.nx), np.float64)
    lib().orc_exact_bp(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_double), len(ang),
                       _ptr(p, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def exact_count(vol, det, angles, per_ray=False):
    """Sum over rays of the number of voxels with nonzero exact length."""
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    counts = np.zeros((len(ang), det.nb, det.na), np.int32) if per_ray else None
    total = lib().orc_exact_count(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_double), len(ang),
                                  _ptr(counts, ctypes.c_int32) if per_ray else None)
    return (int(total), counts) if per_ray else int(total)


def exact_ray(vol, img_zyx, S, D):
    img = np.ascontiguousarray(img_zyx, dtype=np.float64)
    s = np.ascontiguousarray(S, np.float64)
    d = np.ascontiguousarray(D, np.float64)
    return float(lib().orc_exact_ray(ctypes.byref(vol), _ptr(img, ctypes.c_double),
                                     _ptr(s, ctypes.c_double), _ptr(d, ctypes.c_double)))


def exact_ray_count(vol, S, D):
    s = np.ascontiguousarray(S, np.float64)
    d = np.ascontiguousarray(D, np.float64)
    tot = ctypes.c_double(0.0)
    n = lib().orc_exact_ray_count(ctypes.byref(vol), _ptr(s, ctypes.c_double), _ptr(d, ctypes.c_double),
                                  ctypes.byref(tot))
    return int(n), float(tot.value)


def ray_points(det, angle, ia, ib):
    S = np.zeros(3)
    D = np.zeros(3)
    lib().orc_ray_points(ctypes.byref(det), float(angle), int(ia), int(ib),
                         _ptr(S, ctypes.c_double), _ptr(D, ctypes.c_double))
    return S, D


# --------------------------------------------------------------------------
# FP32 restatement of the reference slab walk
# --------------------------------------------------------------------------
def ref_fwd(vol, det, angles, img_zyx, flags=REF_FWD_FLAGS, v0=0, vstep=1):
    """Reference ProjectFlat/ProjectCyli restated (siddon.cu.cc:12-390)."""
    img = np.ascontiguousarray(img_zyx, dtype=np.float32)
    ang = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((len(ang), det.nb, det.na), np.float32)
    lib().orc_ref_fwd(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_float), len(ang),
                      _ptr(img, ctypes.c_float), _ptr(out, ctypes.c_float), flags, v0, vstep)
    return out


def ref_bp(vol, det, angles, pv, flags=REF_BP_FLAGS, v0=0, vstep=1):
    """Reference BackProjectFlat/BackProjectCyli restated (siddon.cu.cc:450-832),
    flat-detector sinogram indexing for both modes (Q5 not reproduced)."""
    p = np.ascontiguousarray(pv, dtype=np.float32)
    ang = np.ascontiguousarray(angles, dtype=np.float32)
    out = np.zeros((vol.nz, vol.ny, vol.nx), np.float32)
    lib().orc_ref_bp(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_float), len(ang),
                     _ptr(p, ctypes.c_float), _ptr(out, ctypes.c_float), flags, v0, vstep)
    return out


def ref_fwd64(vol, det, angles, img_zyx, flags=REF_FWD_FLAGS):
    """The reference slab rule evaluated in FP64 (same restatement, no FP32
    rounding): the target for the reference-mode GPU kernels, which follow the
    rule with their own FP32 arithmetic.  Angles are taken as given (pass the
    float32 angles the GPU sees, widened)."""
    img = np.ascontiguousarray(img_zyx, dtype=np.float64)
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros((len(ang), det.nb, det.na), np.float64)
    lib().orc_ref_fwd_f64(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_double), len(ang),
                          _ptr(img, ctypes.c_double), _ptr(out, ctypes.c_double), flags, 0, 1)
    return out


def ref_bp64(vol, det, angles, pv, flags=REF_BP_FLAGS):
    p = np.ascontiguousarray(pv, dtype=np.float64)
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros((vol.nz, vol.ny, vol.nx), np.float64)
    lib().orc_ref_bp_f64(ctypes.byref(vol), ctypes.byref(det), _ptr(ang, ctypes.c_double), len(ang),
                         _ptr(p, ctypes.c_double), _ptr(out, ctypes.c_double), flags, 0, 1)
    return out


# --------------------------------------------------------------------------
# Trilinear warp (cpp/deform/deform.cu.cc:13-46)
# --------------------------------------------------------------------------
def _warp_corners(shape_zyx, mx, my, mz, boundary):
    """Corner indices/weights/validity of the pull-trilinear warp.

    boundary 'ref': the voxel is written only when every base index satisfies
    1 <= floor(p) and floor(p)+1 <= n-1 (deform.cu.cc:31-33), else 0 (the output
    is zero-filled by nefct/cc/deform/deform.cc:94).
    boundary 'zero': every in-range corner contributes (zero padding).
    """
    nz, ny, nx = shape_zyx
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    # the sample position is formed in float32 as the reference does
    # (deform.cu.cc:25-36: floor(ix + mx[id]) and ix + mx[id] - xi); weights in FP64
    px = (ix.astype(np.float32) + np.asarray(mx, np.float32)).astype(np.float64)
    py = (iy.astype(np.float32) + np.asarray(my, np.float32)).astype(np.float64)
    pz = (iz.astype(np.float32) + np.asarray(mz, np.float32)).astype(np.float64)
    x0 = np.floor(px).astype(np.int64)
    y0 = np.floor(py).astype(np.int64)
    z0 = np.floor(pz).astype(np.int64)
    fx, fy, fz = px - x0, py - y0, pz - z0
    if boundary == "ref":
        ok = (x0 > 0) & (x0 + 1 < nx) & (y0 > 0) & (y0 + 1 < ny) & (z0 > 0) & (z0 + 1 < nz)
    corners = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                cx, cy, cz = x0 + dx, y0 + dy, z0 + dz
                wx = fx if dx else 1 - fx
                wy = fy if dy else 1 - fy
                wz = fz if dz else 1 - fz
                if boundary == "ref":
                    valid = ok
                else:
                    valid = (cx >= 0) & (cx < nx) & (cy >= 0) & (cy < ny) & (cz >= 0) & (cz < nz)
                lin = np.where(valid, cx + nx * (cy + ny * cz), 0)
                corners.append((lin, wx, wy, wz, valid, dx, dy, dz))
    return corners


def deform(img_zyx, mx, my, mz, boundary="ref"):
    """Pull-trilinear warp out[i] = img(i + m(i)), m in voxel units, FP64."""
    img = np.asarray(img_zyx, np.float64)
    flat = img.ravel()
    out = np.zeros(img.shape, np.float64)
    for lin, wx, wy, wz, valid, *_ in _warp_corners(img.shape, mx, my, mz, boundary):
        out += np.where(valid, flat[lin] * wx * wy * wz, 0.0)
    return out


def deform_adjoint(grad_zyx, mx, my, mz, boundary="ref"):
    """Exact transpose of deform with respect to the image (scatter)."""
    g = np.asarray(grad_zyx, np.float64)
    out = np.zeros(g.size, np.float64)
    for lin, wx, wy, wz, valid, *_ in _warp_corners(g.shape, mx, my, mz, boundary):
        np.add.at(out, lin[valid], (g * wx * wy * wz)[valid])
    return out.reshape(g.shape)


def deform_dvf_grad(img_zyx, grad_zyx, mx, my, mz, boundary="ref"):
    """d<grad, deform(img, m)>/d m for each of mx, my, mz (voxel units)."""
    img = np.asarray(img_zyx, np.float64)
    flat = img.ravel()
    g = np.asarray(grad_zyx, np.float64)
    gx = np.zeros(img.shape)
    gy = np.zeros(img.shape)
    gz = np.zeros(img.shape)
    for lin, wx, wy, wz, valid, dx, dy, dz in _warp_corners(img.shape, mx, my, mz, boundary):
        v = np.where(valid, flat[lin], 0.0)
        sx = 1.0 if dx else -1.0
        sy = 1.0 if dy else -1.0
        sz = 1.0 if dz else -1.0
        gx += v * sx * wy * wz
        gy += v * wx * sy * wz
        gz += v * wx * wy * sz
    return gx * g, gy * g, gz * g


# --------------------------------------------------------------------------
# FDK pieces (mat_codes/filtering.m, ParkerWeight.m, backprojection.m)
# --------------------------------------------------------------------------
def detector_coords(n, d, off_frac):
    """Ray coordinates of detector cells, siddon-op convention: (off + i) d - n d / 2."""
    return (off_frac + np.arange(n)) * d - n * d / 2.0


def cosine_weight(det):
    """filtering.m:10: SDD / sqrt(SDD^2 + u^2 + v^2), shape (nb, na); flat detector."""
    u = detector_coords(det.na, det.da, det.ai)
    v = detector_coords(det.nb, det.db, det.bi)
    uu, vv = np.meshgrid(u, v)
    return det.sad / np.sqrt(det.sad ** 2 + uu ** 2 + vv ** 2)


def parker_weights(nv, dang_rad, det, direction=1):
    """ParkerWeight.m:5-27, shape (nv, na).  View i (0-based) sits at
    beta_i = ((N180 - nv)/2 + i + 1) * dang (degrees), N180 = 180/dang."""
    u = detector_coords(det.na, det.da, det.ai)
    as_deg = np.degrees(np.arctan(u / det.sad))
    fan = abs(as_deg[-1] - as_deg[0]) / 2.0
    dang = np.degrees(dang_rad)
    n180 = 180.0 / dang
    w = np.ones((nv, det.na))
    for i in range(nv):
        p = ((n180 - nv) / 2.0 + (i + 1)) / n180 * 180.0 + direction * as_deg
        w[i] = parker_w(p, fan)
    return w


def parker_w(p, fan):
    """Weight as a function of p = beta + gamma (degrees), ParkerWeight.m:11-24."""
    p = np.asarray(p, np.float64)
    w = np.ones_like(p)
    r1 = (p <= fan / 2) & (p > -fan / 2)
    w[r1] = np.maximum(np.sin(np.pi / 2 * (fan / 2 + p[r1]) / fan), 0) ** 2
    r2 = (p >= 180 - fan / 2) & (p < 180 + fan / 2)
    w[r2] = np.maximum(np.cos(np.pi / 2 * (fan / 2 + p[r2] - 180) / fan), 0) ** 2
    r3 = (p < -fan / 2) | (p > 180 + fan / 2)
    w[r3] = 0.0
    return w


def ramp_response(nu, window="ram-lak", d=1.0):
    """filtering.m:20-24,54-88: frequency response of length max(64, 2^nextpow2(2 nu))."""
    n = max(64, int(2 ** np.ceil(np.log2(2 * nu))))
    nn = np.arange(-(n // 2), n // 2)
    h = np.zeros(n)
    h[n // 2] = 0.25
    odd = (nn % 2) == 1
    h[odd] = -1.0 / (np.pi * nn[odd]) ** 2
    fk = np.abs(np.fft.fft(h)) * 2
    filt = fk[: n // 2 + 1].copy()
    w = 2 * np.pi * np.arange(filt.size) / n
    wl = window.lower()
    if wl == "ram-lak":
        pass
    elif wl == "shepp-logan":
        filt[1:] *= np.sin(w[1:] / (2 * d)) / (w[1:] / (2 * d))
    elif wl == "cosine":
        filt[1:] *= np.cos(w[1:] / (2 * d))
    elif wl == "hamming":
        filt[1:] *= 0.54 + 0.46 * np.cos(w[1:] / d)
    elif wl == "hann":
        filt[1:] *= (1 + np.cos(w[1:] / d)) / 2
    else:
        raise ValueError(window)
    filt[w > np.pi * d] = 0
    return np.concatenate([filt, filt[-2:0:-1]])


def ramp_filter_rows(proj, window="ram-lak"):
    """filtering.m:29-41 for one or many projections: zero-pad each detector row
    (length nu = last axis) into the middle of a length-n buffer, multiply its FFT
    by the response, take the real inverse and crop the same range (no scaling)."""
    p = np.asarray(proj, np.float64)
    nu = p.shape[-1]
    filt = ramp_response(nu, window)
    n = filt.size
    buf = np.zeros(p.shape[:-1] + (n,))
    s = n // 2 - nu // 2
    buf[..., s:s + nu] = p
    out = np.real(np.fft.ifft(np.fft.fft(buf, axis=-1) * filt, axis=-1))
    return out[..., s:s + nu]


def fdk_scale(det, nv, dang_rad, parker):
    """filtering.m:43-47 scaling: /2/du * (dbeta/2) * (SDD/SOD), dbeta = 2 pi / nv
    for a full scan, or 2*dang (Parker short scan, the 2 pi/(180/dang) term)."""
    sdd, sod = det.sad, det.sid
    dbeta = (2 * np.pi * np.degrees(dang_rad) / 180.0) if parker else (2 * np.pi / nv)
    return 1.0 / 2.0 / det.da * dbeta / 2.0 * (sdd / sod)


def fdk_backproject(vol, det, angles, q):
    """Voxel-driven weighted backprojection (backprojection.m:11-25) in the
    siddon-op geometry: for voxel centre P and view t, U = SOD + x cos t + y sin t,
    detector coordinates a = (-x sin t + y cos t) SDD/U, b = z SDD/U, bilinear
    sample between cell centres (outside -> 0, backprojection.m:28), weight
    (SOD/U)^2.  Returns [nz, ny, nx] float64 (unscaled sum over views)."""
    nx, ny, nz = vol.nx, vol.ny, vol.nz
    vx, vy, vz = vol.sx / nx, vol.sy / ny, vol.sz / nz
    x = vol.cx - vol.sx / 2 + (np.arange(nx) + 0.5) * vx
    y = vol.cy - vol.sy / 2 + (np.arange(ny) + 0.5) * vy
    z = vol.cz - vol.sz / 2 + (np.arange(nz) + 0.5) * vz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    out = np.zeros((nz, ny, nx))
    sod, sdd = det.sid, det.sad
    for iv, t in enumerate(np.asarray(angles, np.float64)):
        c, s = np.cos(t), np.sin(t)
        U = sod + X * c + Y * s
        mag = sdd / U
        a = (-X * s + Y * c) * mag
        b = Z * mag
        fa = (a + det.na * det.da / 2) / det.da - det.ai
        fb = (b + det.nb * det.db / 2) / det.db - det.bi
        ok = (fa >= 0) & (fa <= det.na - 1) & (fb >= 0) & (fb <= det.nb - 1)
        ia = np.clip(np.floor(fa).astype(np.int64), 0, max(det.na - 2, 0))
        ib = np.clip(np.floor(fb).astype(np.int64), 0, max(det.nb - 2, 0))
        ta, tb = fa - ia, fb - ib
        qq = q[iv]
        v = (qq[ib, ia] * (1 - ta) * (1 - tb) + qq[ib, ia + 1] * ta * (1 - tb)
             + qq[ib + 1, ia] * (1 - ta) * tb + qq[ib + 1, ia + 1] * ta * tb)
        out += np.where(ok, v * (sod / U) ** 2, 0.0)
    return out


# --------------------------------------------------------------------------
# anisotropic TV (nefct/functions/total_variation.py:19-35, :125-149)
# --------------------------------------------------------------------------
def tv_forward(img):
    """Forward differences along each axis with a zero first slice."""
    x = np.asarray(img, np.float64)
    out = np.zeros((x.ndim,) + x.shape)
    for ax in range(x.ndim):
        sl_hi = [slice(None)] * x.ndim
        sl_lo = [slice(None)] * x.ndim
        sl_hi[ax] = slice(1, None)
        sl_lo[ax] = slice(None, -1)
        out[ax][tuple(sl_hi)] = x[tuple(sl_hi)] - x[tuple(sl_lo)]
    return out


def tv_adjoint(d):
    """Transpose of tv_forward (InvertTotalVari)."""
    d = np.asarray(d, np.float64)
    out = np.zeros(d.shape[1:])
    nd = d.shape[0]
    for ax in range(nd):
        sl_hi = [slice(None)] * nd
        sl_lo = [slice(None)] * nd
        sl_hi[ax] = slice(1, None)
        sl_lo[ax] = slice(None, -1)
        out[tuple(sl_lo)] -= d[ax][tuple(sl_hi)]
        out[tuple(sl_hi)] += d[ax][tuple(sl_hi)]
    return out


# --------------------------------------------------------------------------
# Voxel-driven unweighted backprojection (nefct/cc/pixel_3d/pixel.cu.cc:12-124)
# and pixel-driven FDK footprint backprojection
# (nefct/cc/distance_3d/dist_3d_back_fdk.cu.cc:17-255), C restatements in FP64
# --------------------------------------------------------------------------
def pixel_bp(cyl, grid, center_low, size, angles, SID, SAD, na, da, ai, nb, db, bi, pv):
    """grid (nx, ny, nz); center_low = the reference op's ``center`` (voxel
    centres at (i + 0.5) size / n + center); pv [nv, nb, na] -> [nz, ny, nx]."""
    nx, ny, nz = (int(v) for v in grid)
    ang = np.ascontiguousarray(angles, np.float64)
    pv = np.ascontiguousarray(pv, np.float64)
    assert pv.shape == (ang.size, nb, na)
    out = np.empty((nz, ny, nx), np.float64)
    g = np.array([nx, ny, nz], np.int32)
    c = np.ascontiguousarray(center_low, np.float64)
    z = np.ascontiguousarray(size, np.float64)
    lib().orc_pixel_bp(int(bool(cyl)), _ptr(g, ctypes.c_int), _ptr(c, ctypes.c_double), _ptr(z, ctypes.c_double),
                       _ptr(ang, ctypes.c_double), ang.size, float(SID), float(SAD), int(na), float(da), float(ai),
                       int(nb), float(db), float(bi), _ptr(pv, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def fdk_fp_bp(cyl, grid, angles, offsets, SID, SAD, na, da, ai, nb, db, bi, pv):
    """Footprint FDK backprojection in voxel units, naming as in
    dist_3d_back_fdk.cu.cc (SAD = source-axis, SID = source-detector);
    pv [nv, nb, na] -> [nz, ny, nx]."""
    nx, ny, nz = (int(v) for v in grid)
    ang = np.ascontiguousarray(angles, np.float64)
    pv = np.ascontiguousarray(pv, np.float64)
    assert pv.shape == (ang.size, nb, na)
    off = None if offsets is None else np.ascontiguousarray(offsets, np.float64)
    out = np.empty((nz, ny, nx), np.float64)
    lib().orc_fdk_fp_bp(int(bool(cyl)), nx, ny, nz, _ptr(ang, ctypes.c_double),
                        None if off is None else _ptr(off, ctypes.c_double), ang.size, float(SID), float(SAD),
                        int(na), float(da), float(ai), int(nb), float(db), float(bi), _ptr(pv, ctypes.c_double),
                        _ptr(out, ctypes.c_double))
    return out


# --------------------------------------------------------------------------
# Texture-sampled warps (nefct/cc/deform/deform_tex.cu.cc)
# --------------------------------------------------------------------------
def tex_sample(vol_zyx, u, v, w):
    """Trilinear sample with the semantics of the reference's 3-D texture
    (deform_tex.cu.cc:60-69: unnormalised coordinates, linear filtering,
    clamp addressing): along each axis the sample coordinate u covers texel i
    on [i, i + 1), so the value is (1 - a) T[clamp(i0)] + a T[clamp(i0 + 1)]
    with i0 = floor(u - 1/2), a = u - 1/2 - i0 and clamp to [0, n - 1].
    Weights are exact (FP64); the hardware's 1/256 quantisation of a is not
    modelled.  u, v, w are arrays of one shape (x, y, z coordinates)."""
    vol = np.asarray(vol_zyx, np.float64)
    nz, ny, nx = vol.shape
    out = np.zeros(np.shape(u), np.float64)
    axes = []
    for c, n in ((u, nx), (v, ny), (w, nz)):
        xb = np.asarray(c, np.float64) - 0.5
        i0 = np.floor(xb)
        a = xb - i0
        i0 = i0.astype(np.int64)
        axes.append((np.clip(i0, 0, n - 1), np.clip(i0 + 1, 0, n - 1), a))
    (x0, x1, ax), (y0, y1, ay), (z0, z1, az) = axes
    for zi, wz in ((z0, 1 - az), (z1, az)):
        for yi, wy in ((y0, 1 - ay), (y1, ay)):
            for xi, wx in ((x0, 1 - ax), (x1, ax)):
                out += vol[zi, yi, xi] * wx * wy * wz
    return out


def _tex_grid(shape_zyx):
    nz, ny, nx = shape_zyx
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ix, iy, iz


def _f32(*xs):
    """Sum in float32 left to right, as the reference forms its coordinates."""
    acc = np.float32(0) + np.asarray(xs[0], np.float32)
    for x in xs[1:]:
        acc = (acc + np.asarray(x, np.float32)).astype(np.float32)
    return acc.astype(np.float64)


def deform_tex(img_zyx, mx, my, mz):
    """DeformKernel, deform_tex.cu.cc:15-26: out[i] = T(i + m(i) + 1/2)."""
    ix, iy, iz = _tex_grid(np.shape(img_zyx))
    return tex_sample(img_zyx, _f32(ix, mx, 0.5), _f32(iy, my, 0.5), _f32(iz, mz, 0.5))


def invert_dvf(mx, my, mz, iters=30):
    """DeformInvertKernel, deform_tex.cu.cc:78-100: per voxel i the fixed-point
    iteration x = -mx(i + d), y = -my(i + d), z = -mz(i + d) with d = (x, y, z)
    updated component by component (each sample uses the components already
    updated in this iteration), starting from d = 0, 30 iterations.  FP64."""
    mx, my, mz = (np.asarray(a, np.float64) for a in (mx, my, mz))
    ix, iy, iz = _tex_grid(mx.shape)
    x = np.zeros(mx.shape)
    y = np.zeros(mx.shape)
    z = np.zeros(mx.shape)
    for _ in range(iters):
        x = -tex_sample(mx, x + ix + 0.5, y + iy + 0.5, z + iz + 0.5)
        y = -tex_sample(my, x + ix + 0.5, y + iy + 0.5, z + iz + 0.5)
        z = -tex_sample(mz, x + ix + 0.5, y + iy + 0.5, z + iz + 0.5)
    return x, y, z


def deform_invert_tex(img_zyx, mx, my, mz, iters=30):
    """deform_invert_tex, deform_tex.cu.cc:170-226: warp by the inverted field."""
    x, y, z = invert_dvf(mx, my, mz, iters)
    return deform_tex(img_zyx, x, y, z)


def shift_tex(img_zyx, sx, sy, sz):
    """ShiftKernel, deform_tex.cu.cc:228-239: out[i] = T(i - s + 1/2)."""
    ix, iy, iz = _tex_grid(np.shape(img_zyx))
    return tex_sample(img_zyx, _f32(ix, -np.float32(sx), 0.5), _f32(iy, -np.float32(sy), 0.5),
                      _f32(iz, -np.float32(sz), 0.5))


def shift_z_tex(img_zyx, sz):
    """ShiftZKernel, deform_tex.cu.cc:291-302: shift along z only."""
    ix, iy, iz = _tex_grid(np.shape(img_zyx))
    return tex_sample(img_zyx, ix + 0.5, iy + 0.5, _f32(iz, -np.float32(sz), 0.5))


def rotate_tex(img_zyx, cos_phi, sin_phi):
    """RotateKernel, deform_tex.cu.cc:355-370: in-plane rotation about the
    integer centre (nx/2, ny/2) (integer division), sampled at
    (x', y', iz + 1/2) with x' = (ix + 1/2 - nx/2) cos - (iy + 1/2 - ny/2) sin + nx/2,
    y' = (ix + 1/2 - nx/2) sin + (iy + 1/2 - ny/2) cos + ny/2 (float32, as the
    reference evaluates them)."""
    nz, ny, nx = np.shape(img_zyx)
    ix, iy, iz = _tex_grid((nz, ny, nx))
    c, s = np.float32(cos_phi), np.float32(sin_phi)
    ax = (ix.astype(np.float32) + np.float32(0.5) - np.float32(nx // 2)).astype(np.float32)
    ay = (iy.astype(np.float32) + np.float32(0.5) - np.float32(ny // 2)).astype(np.float32)
    xp = ((ax * c).astype(np.float32) - (ay * s).astype(np.float32)).astype(np.float32) + np.float32(nx // 2)
    yp = ((ax * s).astype(np.float32) + (ay * c).astype(np.float32)).astype(np.float32) + np.float32(ny // 2)
    return tex_sample(img_zyx, xp.astype(np.float64), yp.astype(np.float64), iz + 0.5)
