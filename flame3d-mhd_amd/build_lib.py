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


H5_LIB = os.path.join(HERE, "libf3d_h5.so")
H5_PREFIX = os.environ.get("F3D_HDF5_PREFIX", "/opt/conda")


def build_h5(verbose=False):
    """libf3d_h5.so: the on-disk formats (HDF5 C library, host code, g++).  Returns None when the HDF5
    headers are absent (the GPU library does not depend on it)."""
    inc, libdir = os.path.join(H5_PREFIX, "include"), os.path.join(H5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        sys.stderr.write(f"[build] no hdf5.h under {inc}: libf3d_h5.so not built\n")
        return None
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall", "-Wextra",
           f"-I{inc}", f"-I{os.path.join(os.path.dirname(HERE), 'include')}", os.path.join(CSRC, "f3d_h5io.cpp"),
           "-o", tmp, f"-L{libdir}", "-lhdf5", f"-Wl,-rpath,{libdir}", "-static-libstdc++", "-static-libgcc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or r.returncode:
        sys.stderr.write(f"[build] {' '.join(cmd)}\n{r.stdout}{r.stderr}")
    if r.returncode:
        raise RuntimeError("build of libf3d_h5.so failed")
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_h5(verbose="-v" in sys.argv))
