"""Builds libf3d_hip.so (the HIP/gfx950 C-ABI library) in-tree with hipcc.

Run as a script or call build().  Every call recompiles; nothing is cached outside the tree.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libf3d_hip.so")
SOURCES = ["f3d_capi.hip", "f3d_kernels.hip", "f3d_stage.hip", "f3d_fused.hip", "f3d_fvm.hip", "f3d_adjoint.hip",
           "f3d_timing.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 (v_pk_*_f32) generation disabled for the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, out, r.returncode, r.stdout + r.stderr, cmd


def build(verbose=False, jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = []
    for src, out, rc, log, cmd in results:
        if verbose or rc:
            sys.stderr.write(f"[build] {' '.join(cmd)}\n{log}")
        if rc:
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libf3d_hip.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
