"""ctypes binding of libnefct_hip.so (the C ABI declared in include/nefct_hip.h).

The library is built in-tree by ``ct-nef_amd/build.py``.  There is no CPU
fallback: if the library is missing or no HIP device is present, every
operator raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnefct_hip.so")

# flags (include/nefct_hip.h)
F_INV_L2 = 1
F_REF_SLAB = 2
F_SHORT_SKIP = 4
F_TIE_GT = 8
F_DETERMINISTIC = 16
F_ADJ_WINDOW = 32
F_ADJ_FLOAT_ATOMIC = 0x1000
F_TILE_MASK = 0x700
F_TILE_32x32x16 = 0x000
F_TILE_32x16x8 = 0x100
F_TILE_32x32x8 = 0x200
F_TILE_64x32x8 = 0x300
F_TILE_16x16x32 = 0x400
F_TILE_64x16x16 = 0x500
F_LANE_STRIDE2 = 0x800
F_FWD_TILE_MASK = 0x6000
F_FWD_TILE_16x16 = 0x0000
F_FWD_TILE_32x8 = 0x2000
F_FWD_TILE_8x32 = 0x4000
F_FWD_TILE_64x4 = 0x6000
F_FWD_VG_SHIFT = 15
F_FWD_VG_MASK = 7 << F_FWD_VG_SHIFT


def fwd_vg(log2_views: int) -> int:
    """Forward view-group field for 2^log2_views views (1..6); 0 = default rule."""
    if not 0 <= log2_views <= 6:
        raise ValueError("view group log2 must be in 0..6")
    return log2_views << F_FWD_VG_SHIFT
REF_FORWARD = F_INV_L2 | F_REF_SLAB
REF_ADJOINT = F_REF_SLAB | F_SHORT_SKIP | F_TIE_GT

_c_int3 = ctypes.c_int * 3
_c_float3 = ctypes.c_float * 3
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_u = ctypes.c_uint
_sz = ctypes.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "nefct_status_string": [_i],
    "nefct_siddon_workspace_size": [_vp, _i],
    "nefct_project_flat": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _u, _vp, _sz, _vp],
    "nefct_project_cyli": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _u, _vp, _sz, _vp],
    "nefct_backproject_flat": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _u, _vp, _sz, _vp],
    "nefct_backproject_flat_noalloc": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _u, _vp],
    "nefct_backproject_cyli": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _u, _vp, _sz, _vp],
    "nefct_project_helical": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _vp, _u, _vp,
                              _sz, _vp],
    "nefct_backproject_helical": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _vp, _u,
                                  _vp, _sz, _vp],
    "nefct_siddon_count": [_vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _vp, _u, _vp, _sz, _vp],
    "nefct_deform": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "nefct_deform_adjoint": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "nefct_deform_dvf_grad": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp],
    "nefct_deform_tex": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "nefct_invert_dvf": [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "nefct_deform_invert_tex": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp],
    "nefct_shift_tex": [_vp, _vp, _f, _f, _f, _vp, _vp],
    "nefct_shift_z_tex": [_vp, _vp, _f, _vp, _vp],
    "nefct_rotate_tex": [_vp, _vp, _f, _f, _vp, _vp],
    "nefct_fdk_preweight": [_vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _i, _f, _i, _f, _vp],
    "nefct_fdk_filter": [_vp, _vp, _i, _i, _vp, _vp],
    "nefct_fdk_workspace_size": [_i],
    "nefct_fdk_backproject": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _vp, _sz, _vp],
    "nefct_fdk_backproject_mapped": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _i, _vp,
                                     _sz, _vp],
    "nefct_pixel_backproject_flat": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _vp],
    "nefct_pixel_backproject_cyli": [_vp, _vp, _vp, _vp, _vp, _i, _f, _f, _i, _f, _f, _i, _f, _f, _vp, _vp],
    "nefct_fdk_footprint_backproject_flat": [_vp, _vp, _vp, _vp, _i, _f, _f, _f, _f, _i, _f, _f, _i, _vp, _vp],
    "nefct_fdk_footprint_backproject_cyli": [_vp, _vp, _vp, _vp, _i, _f, _f, _f, _f, _i, _f, _f, _i, _vp, _vp],
    "nefct_tv_smooth": [_vp, _vp, _f, _f, _vp, _vp, _vp],
    "nefct_sq_norm": [_vp, _sz, ctypes.c_double, _vp, _vp],
    "nefct_tv_descent_step": [_vp, _vp, _vp, _f, _f, _f, _vp, _vp],
    "nefct_batch_plan": [_vp, _i, _u, _vp, _vp, _vp],
    "nefct_batch_project": [_vp, _vp, _i, ctypes.c_longlong, _u, _vp, _sz, _vp],
    "nefct_batch_backproject": [_vp, _vp, _i, ctypes.c_longlong, _u, _vp, _sz, _vp],
}
_RESTYPE = {
    "nefct_status_string": ctypes.c_char_p,
    "nefct_siddon_workspace_size": ctypes.c_size_t,
    "nefct_fdk_workspace_size": ctypes.c_size_t,
}

_lib = None


class NefctError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NefctError(f"{LIB_PATH} is missing: run ct-nef_amd/build.py (hipcc, gfx950) first")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = lib
    return _lib


def check(status: int, what: str):
    if status != 0:
        msg = load().nefct_status_string(int(status)).decode()
        raise NefctError(f"{what}: {msg} (status {status})")


def int3(v):
    return _c_int3(*(int(x) for x in v))


def float3(v):
    return _c_float3(*(float(x) for x in v))
