"""Build the gfx950 HIP library libnefct_hip.so in-tree (hipcc, no JIT cache).

Every HIP source under csrc/ is compiled for gfx950 only (one object per
source, in parallel) and linked into one shared library placed next to the
Python package (nefct_hip/libnefct_hip.so), so it travels with the source
tree.  Packed FP32 instructions are disabled (-target-feature
-packed-fp32-ops) on purpose.

siddon.hip's device code is compiled with -fno-honor-nans (device side only,
host code keeps IEEE semantics): (int)floorf(x) in the walk loops then
becomes a single v_cvt_flr_i32_f32.  That file tests non-finite data on the
bits, so its NaN behaviour does not depend on the flag.
"""
import glob
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "nefct_hip", "libnefct_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]
PER_FILE = {
    "siddon.hip": ["-Xarch_device", "-fno-honor-nans"],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (res.stderr or "").splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + err)


def build(verbose=False):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    inc = ["-I", os.path.join(HERE, "..", "include")]
    with tempfile.TemporaryDirectory(prefix="nefct_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s) + ".o") for s in srcs]
        cmds = [[HIPCC] + FLAGS + PER_FILE.get(os.path.basename(s), []) + inc + ["-c", s, "-o", o]
                for s, o in zip(srcs, objs)]
        with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1, 8)) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        tmp = OUT + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
