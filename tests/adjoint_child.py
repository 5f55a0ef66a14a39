"""Child process of test_gpu_adjoint.py: runs the library's checkpointed adjoint on the inputs held in
one .npz under the environment switches its parent set (the library reads them once per process) and
writes the objective and the gradients to another .npz.  Usage: adjoint_child.py IN.npz OUT.npz"""
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

    nsteps, every = int(z["nsteps"]), int(z["every"])
    obj = torch.zeros(nsteps, dtype=torch.float64, device="cuda:0")
    lamU0 = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    lin = torch.zeros(z["inflow"].shape, dtype=torch.float64, device="cuda:0")
    lbc = torch.zeros(7, dtype=torch.float64, device="cuda:0")
    m = [d(z[f"m{i}"]) for i in range(9)]
    ctx.adjoint(nsteps, every, d(z["U0"]), m, d(z["J"]), float(z["dt"]), d(z["W"]), obj, lamU0, d(z["inflow"]), lin,
                lbc)
    torch.cuda.synchronize()
    np.savez(dst, obj=obj.cpu().numpy(), lamU0=lamU0.cpu().numpy(), lin=lin.cpu().numpy(), lbc=lbc.cpu().numpy())


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
