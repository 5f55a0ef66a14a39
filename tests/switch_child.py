"""Child process of test_gpu_switches.py: runs the library's RK3 time step on the inputs held in one
.npz under the environment switches its parent set (the library reads them once per process) and
writes U and Q after the steps to another .npz.  Usage: switch_child.py IN.npz OUT.npz"""
import importlib
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main(src, dst):
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    f3d = importlib.import_module("flame3d-mhd_amd")
    z = np.load(src)
    p = f3d.F3DParams.from_buffer_copy(z["params"].tobytes())
    ctx = f3d.Flame3D(p)

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")

    out = {}
    for tag in ("f64", "f32b"):
        U = d(z[f"{tag}_U"])
        Un = torch.zeros_like(U)
        Q = d(z[f"{tag}_Q"])
        m = [d(z[f"{tag}_m{i}"]) for i in range(9)]
        ctx.time_step(int(z["nsteps"]), U, Un, Q, m, d(z["J"]), float(z[f"{tag}_dt"]), d(z[f"{tag}_inflow"]))
        torch.cuda.synchronize()
        out[f"{tag}_U"] = U.cpu().numpy()
        out[f"{tag}_Q"] = Q.cpu().numpy()
    np.savez(dst, **out)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
