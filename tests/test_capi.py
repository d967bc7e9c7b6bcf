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


def test_problem_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of nefct_problem has the C layout (compiled with gcc)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from nefct_hip.batch import Problem
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "nefct_hip.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu\\n", sizeof(nefct_problem),'
                   ' offsetof(nefct_problem, tile_begin), offsetof(nefct_problem, blocks),'
                   ' offsetof(nefct_problem, SID), offsetof(nefct_problem, ws_offset)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [ctypes.sizeof(Problem), Problem.tile_begin.offset, Problem.blocks.offset, Problem.SID.offset,
                   Problem.ws_offset.offset]
