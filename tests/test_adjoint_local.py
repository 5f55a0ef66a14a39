"""CPU check of the adjoint's local reverse-mode pieces (split, WENO7/WENO5/NND faces with the
characteristic frame, viscous cell, c2Prim): the library's own host+device functions, compiled
host-only, against dual-number forward mode of its forward pieces on random inputs
(tests/native/adj_local_check.cpp).  No GPU needed."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_local_vjp_pieces_match_forward_mode(tmp_path):
    exe = tmp_path / "adj_local_check"
    src = os.path.join(HERE, "native", "adj_local_check.cpp")
    r = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
