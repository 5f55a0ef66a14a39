"""From the reference's files to the GPU path: mesh.h5 -> f3d_metrics_f64 on the device, metrics.h5 +
chk-*.h5 -> a restarted time step, both against the same computation from the arrays in memory."""
import importlib
import os

import numpy as np
import pytest

from casegen import f3d, ramp_case, relerr
from test_gpu_parity import dev, host

h5 = importlib.import_module("flame3d-mhd_amd.h5io")
pytestmark = pytest.mark.gpu
NG = 4


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def test_mesh_file_to_device_metrics(torch_mod, tmp_path):
    torch = torch_mod
    case = ramp_case(24, 16, 8)
    p = case.params
    path = str(tmp_path / "mesh.h5")
    h5.write_mesh(path, p, *case.coords, file_dtype=np.float64)
    x, y, z = h5.read_mesh(path, p)
    I = (slice(NG, -NG),) * 3
    for a, b in zip((x, y, z), case.coords):
        assert np.array_equal(a[I], b[I])
    ctx = f3d.Flame3D(p)
    m = [torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0") for _ in range(9)]
    J = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    ctx.metrics(dev(torch, x), dev(torch, y), dev(torch, z), m, J)
    # away from the boundary the 6th-order stencils read interior points only: identical metrics
    deep = (slice(NG + 3, -NG - 3),) * 3
    for q in range(9):
        assert relerr(host(m[q])[deep], case.metrics[q][deep], axis_vars=False) < 1e-12, q
    assert relerr(host(J)[deep], case.J[deep], axis_vars=False) < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_restart_from_files(torch_mod, tmp_path, dtype):
    torch = torch_mod
    case = ramp_case(24, 16, 8, dtype=dtype)
    p = case.params
    mpath, cpath = str(tmp_path / "metrics.h5"), str(tmp_path / "chk-100.h5")
    h5.write_metrics(mpath, p, [a.astype(np.float64) for a in case.metrics], case.J, file_dtype=np.float64)
    h5.write_checkpoint(cpath, p, case.Q)  # Float32 file, as the reference writes it
    mr, Jr = h5.read_metrics(mpath, p)
    Qr = h5.read_checkpoint(cpath, p, dtype=dtype)
    Q32 = case.Q.astype(np.float32).astype(dtype)
    assert np.array_equal(Qr, Q32)
    ctx = f3d.Flame3D(p)
    U0 = np.zeros(p.cell_shape(5))
    from oracle_lib import Oracle
    Oracle(p).prim2c(U0, Q32)
    dt = 0.2 * Oracle(p).cfl_dt(Q32, case.metrics, case.J, 1.0)
    runs = []
    for Q, m, J in ((Qr, [a.astype(dtype) for a in mr], Jr), (Q32, case.metrics, case.J)):
        Ug, Qg = dev(torch, U0), dev(torch, Q)
        ctx.time_step(2, Ug, torch.zeros_like(Ug), Qg, [dev(torch, np.ascontiguousarray(a)) for a in m],
                      dev(torch, J), dt, dev(torch, case.inflow))
        runs.append((host(Ug), host(Qg)))
    inner = (slice(None),) + (slice(NG, -NG),) * 3
    # same inputs; the heavy-face corrections add with atomics, so the last bits may differ
    assert relerr(runs[0][0][inner], runs[1][0][inner]) < (1e-12 if dtype == np.float64 else 1e-9)
    assert relerr(runs[0][1], runs[1][1]) < (1e-12 if dtype == np.float64 else 1e-6)
