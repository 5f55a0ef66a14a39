"""Child process of test_gpu_sweep_forms.py: evaluates the RHS and `nsteps` time steps (finite-difference,
or finite-volume when the parameters select it) of the inputs held in one .npz under the environment switches its parent set (the library reads them once per process) and
writes R, U, Un and Q to another .npz.  Usage: sweep_child.py IN.npz OUT.npz"""
import importlib
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def run(f3d, torch, z):
    """RHS of Q, then nsteps steps of (U, Q); returns host arrays R, U, Un, Q."""
    p = f3d.F3DParams.from_buffer_copy(z["params"].tobytes())
    ctx = f3d.Flame3D(p)

    def d(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")

    m = [d(z[f"m{i}"]) for i in range(9)]
    J, Q, U = d(z["J"]), d(z["Q"]), d(z["U"])
    inflow = d(z["inflow"]) if z["inflow"].size else None
    R = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    Un = torch.zeros_like(U)
    if p.fvm:  # finite-volume branch: face geometry from the mesh points
        geo = [torch.zeros((4,) + p.face_shape(q)[1:], dtype=Q.dtype, device="cuda:0") for q in range(3)]
        ctx.fvm_geometry(d(z["x"]), d(z["y"]), d(z["z"]), geo)
        ctx.rhs_fvm(R, Q, U, m, J, geo)
        ctx.time_step_fvm(int(z["nsteps"]), U, Un, Q, m, J, geo, float(z["dt"]), inflow)
    else:
        ctx.rhs(R, Q, m, J)
        ctx.time_step(int(z["nsteps"]), U, Un, Q, m, J, float(z["dt"]), inflow)
    torch.cuda.synchronize()
    return dict(R=R.cpu().numpy(), U=U.cpu().numpy(), Un=Un.cpu().numpy(), Q=Q.cpu().numpy())


def main(src, dst):
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    f3d = importlib.import_module("flame3d-mhd_amd")
    np.savez(dst, **run(f3d, torch, np.load(src)))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
