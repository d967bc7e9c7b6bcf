"""The allocation-free adjoint nefct_backproject_flat_noalloc, called through
the C ABI, against nefct_backproject_flat (workspace from the caller) on the
same inputs: two calls in a row on one device, the second with a smaller
volume, detector and view count, so a stale scratch (maximum word, view
table) from the first call would show.  The fixed-point adjoint is bitwise
reproducible, so the bar is bitwise equality; the float-atomic variant is
compared to 1e-6 relative L2.  Sizes beyond the fixed scratch and the window
variant are refused before anything is launched."""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
DEV = "cuda"
_vp = ctypes.c_void_p


def _args(sino, vol, det, ang, out, flags):
    from nefct_hip import _lib
    return [_vp(sino.data_ptr()), _lib.int3(vol.grid), _lib.float3(vol.center), _lib.float3(vol.size),
            _vp(ang.data_ptr()), int(ang.numel()), det.SID, det.SAD, det.na, det.da, det.ai, det.nb, det.db,
            det.bi, _vp(out.data_ptr()), int(flags)]


def _with_workspace(sino, vol, det, ang, flags):
    from nefct_hip import _lib
    L = _lib.load()
    out = torch.full(vol.shape_zyx, float("nan"), device=DEV)
    nbytes = L.nefct_siddon_workspace_size(_lib.int3(vol.grid), int(ang.numel()))
    ws = torch.empty(int(nbytes), dtype=torch.uint8, device=DEV)
    st = L.nefct_backproject_flat(*_args(sino, vol, det, ang, out, flags), _vp(ws.data_ptr()), int(nbytes),
                                  _vp(torch.cuda.current_stream().cuda_stream))
    _lib.check(st, "nefct_backproject_flat")
    return out


def _noalloc(sino, vol, det, ang, flags):
    from nefct_hip import _lib
    L = _lib.load()
    out = torch.full(vol.shape_zyx, float("nan"), device=DEV)
    st = L.nefct_backproject_flat_noalloc(*_args(sino, vol, det, ang, out, flags),
                                          _vp(torch.cuda.current_stream().cuda_stream))
    _lib.check(st, "nefct_backproject_flat_noalloc")
    return out


def _case(side, nz, na, nb, nv, seed, scale):
    import nefct_hip as nh
    vol = nh.VolumeGeometry.cube((side, side, nz), 1.0)
    det = nh.ConeGeometry.flat(300.0, 450.0, na, 1.5 * side / na * 1.1, nb, 1.5 * nz / nb * 1.1)
    ang = torch.tensor((2 * math.pi * np.arange(nv) / nv).astype(np.float32), device=DEV)
    g = torch.Generator(device=DEV).manual_seed(seed)
    sino = scale * torch.randn((nv, nb, na), generator=g, device=DEV)
    return vol, det, ang, sino


@pytest.mark.parametrize("flags", [0, 2 | 4 | 8])   # exact, reference rule
def test_noalloc_equals_workspace_adjoint_twice(flags):
    # first call: larger problem with a large sinogram scale; second: smaller
    # problem, fewer views, scale 1000x smaller (a stale maximum word would
    # change the fixed-point scale and the result's bits)
    for side, nz, na, nb, nv, seed, scale in ((48, 40, 96, 80, 150, 1, 1e3), (24, 20, 40, 32, 37, 2, 1.0)):
        vol, det, ang, sino = _case(side, nz, na, nb, nv, seed, scale)
        want = _with_workspace(sino, vol, det, ang, flags)
        got = _noalloc(sino, vol, det, ang, flags)
        torch.cuda.synchronize()
        assert torch.isfinite(want).all()
        assert torch.equal(got, want), (side, nv, flags)


def test_noalloc_float_atomic_variant():
    from nefct_hip import _lib
    for side, nz, na, nb, nv, seed in ((40, 32, 80, 64, 90, 3), (16, 16, 32, 24, 11, 4)):
        vol, det, ang, sino = _case(side, nz, na, nb, nv, seed, 1.0)
        want = _with_workspace(sino, vol, det, ang, 0)
        got = _noalloc(sino, vol, det, ang, _lib.F_ADJ_FLOAT_ATOMIC)
        rel = (torch.linalg.norm((got - want).double()) / torch.linalg.norm(want.double())).item()
        assert rel < 1e-6, rel


def test_noalloc_refuses_beyond_scratch_and_window():
    from nefct_hip import _lib
    L = _lib.load()
    vol, det, ang, sino = _case(16, 16, 32, 24, 8193, 5, 1.0)
    out = torch.zeros(vol.shape_zyx, device=DEV)
    st = L.nefct_backproject_flat_noalloc(*_args(sino, vol, det, ang, out, 0),
                                          _vp(torch.cuda.current_stream().cuda_stream))
    assert st == -7 and b"8192" in L.nefct_status_string(st)
    vol, det, ang, sino = _case(16, 16, 32, 24, 8, 6, 1.0)
    st = L.nefct_backproject_flat_noalloc(*_args(sino, vol, det, ang, out, _lib.F_ADJ_WINDOW),
                                          _vp(torch.cuda.current_stream().cuda_stream))
    assert st == -5
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0
