"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/nefct_hip.h declares, and the ctypes table covers them.  No compute
call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nefct_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nefct_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("nefct_project_flat", "nefct_project_cyli", "nefct_backproject_flat", "nefct_backproject_cyli"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from nefct_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnefct_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from nefct_hip import _lib
    assert set(_lib.SIGNATURES) == set(declared_symbols())


def test_status_strings_and_workspace_size_without_gpu():
    from nefct_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.load()
    assert L.nefct_status_string(0) == b"ok"
    assert b"workspace" in L.nefct_status_string(-4)
    n = L.nefct_siddon_workspace_size(_lib.int3((512, 512, 256)), 720)
    # two padded copies of the volume + view table
    assert n >= 2 * 4 * 516 * 512 * 260
    assert L.nefct_siddon_workspace_size(_lib.int3((0, 5, 5)), 3) == 0


def test_operators_refuse_cpu_tensors():
    torch = pytest.importorskip("torch")
    import nefct_hip as nh
    vol = nh.VolumeGeometry.cube((8, 8, 8), 1.0)
    det = nh.ConeGeometry.flat(100.0, 150.0, 4, 1.0, 4, 1.0)
    with pytest.raises(nh.NefctError):
        nh.project(torch.zeros(vol.shape_zyx), vol, det, [0.0])
