"""The fused stage's two forms of the x and y faces agree to rounding: the one-march form (x faces formed
in the y march, R written once; the default) against k_sweep_x followed by the y march
(F3D_SWEEP_XY=0, in a child process: the library reads its switches once per process); likewise the
finite-volume branch's k_fvm_sweep_xy against k_fvm_sweep_x + the y march (F3D_FVM_SWEEP_XY=0), whose sums
are the same expressions (the x and y face fluxes are formed apart, then differenced) and agree bit for bit.  The sums are the
same, in the same order; the compiler contracts products into FMAs differently once both splits of a cell
sit in one kernel, so the bar is rounding-level: RHS within 1e-14 (FP64) / 2e-6 (FP32-bulk) of the
per-variable scale of the terms it is summed from, and the state after three time steps (whose last leaves
Un = U^n) within 1e-13 / 1e-6 normwise.  Inputs with faces on WENO7, WENO5 and NND, window and segment
edges that do not divide the grid, both numeric forms."""
import os
import subprocess
import sys

import numpy as np
import pytest

from casegen import f3d, ramp_case, relerr, rhs_scales, scaled_err, tgv_case
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def _case(name, dtype, fvm=0):
    if name == "ramp70x30x9":  # two x windows (56 + 14 columns), y segments that do not divide 30
        return ramp_case(70, 30, 9, dtype=dtype, jump=1.0, spikes=[(5, 3, 2, 20.0), (60, 27, 7, 3.0)],
                         hybrid_phi3=2.5, interior_noise=0.02, fvm=fvm)
    if name == "swept20x12x10":  # one narrow window
        return ramp_case(20, 12, 10, sweep=30, dtype=dtype, jump=0.5, spikes=[(6, 4, 3, 20.0)], hybrid_phi3=2.5,
                         interior_noise=1e-3, fvm=fvm)
    return tgv_case(16, dtype=dtype, fvm=fvm)


def _inputs(tmp_path, case, dtype):
    inflow = case.inflow if case.inflow is not None else np.zeros(0, dtype=dtype)
    arrays = dict(params=np.frombuffer(bytes(case.params), dtype=np.uint8), Q=case.Q, U=case.U(), J=case.J,
                  inflow=inflow, nsteps=np.int64(3),
                  dt=np.float64(0.3 * Oracle(case.params).cfl_dt(case.Q, case.metrics, case.J, 1.0)))
    arrays.update({f"m{i}": a for i, a in enumerate(case.metrics)})
    if case.coords is not None:
        arrays.update(x=case.coords[0], y=case.coords[1], z=case.coords[2])
    src, dst = str(tmp_path / "in.npz"), str(tmp_path / "out.npz")
    np.savez(src, **arrays)
    return src, dst


def _child(src, dst, switch, value="0"):
    env = {k: v for k, v in os.environ.items() if not k.startswith("F3D_")}
    env[switch] = value
    r = subprocess.run([sys.executable, os.path.join(HERE, "sweep_child.py"), src, dst], env=env,
                       capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, f"child exit {r.returncode}\n{r.stderr[-2000:]}"
    return np.load(dst)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32b"])
@pytest.mark.parametrize("name", ["ramp70x30x9", "swept20x12x10", "tgv16"])
def test_sweep_forms_agree(tmp_path, name, dtype):
    import torch
    import sweep_child
    assert torch.cuda.is_available(), "GPU test"
    case = _case(name, dtype)
    inflow = case.inflow if case.inflow is not None else np.zeros(0, dtype=dtype)
    arrays = dict(params=np.frombuffer(bytes(case.params), dtype=np.uint8), Q=case.Q, U=case.U(), J=case.J,
                  inflow=inflow, nsteps=np.int64(3),
                  dt=np.float64(0.3 * Oracle(case.params).cfl_dt(case.Q, case.metrics, case.J, 1.0)))
    arrays.update({f"m{i}": a for i, a in enumerate(case.metrics)})
    src, dst = str(tmp_path / "in.npz"), str(tmp_path / "out.npz")
    np.savez(src, **arrays)
    z = np.load(src)
    mine = sweep_child.run(f3d, torch, z)
    env = {k: v for k, v in os.environ.items() if not k.startswith("F3D_")}
    env["F3D_SWEEP_XY"] = "0"
    print(f"{name} {np.dtype(dtype).name}: child with F3D_SWEEP_XY=0", flush=True)
    r = subprocess.run([sys.executable, os.path.join(HERE, "sweep_child.py"), src, dst], env=env,
                       capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, f"child exit {r.returncode}\n{r.stderr[-2000:]}"
    ref = np.load(dst)
    f64 = dtype == np.float64
    eR = scaled_err(mine["R"], ref["R"], rhs_scales(case))
    print(f"  R: term-scaled |one march - two kernels| = {eR:.2e}")
    assert eR <= (1e-14 if f64 else 2e-6), eR
    for k in ("U", "Un", "Q"):
        e = relerr(mine[k], ref[k])
        print(f"  {k}: normwise {e:.2e}")
        assert e <= (1e-13 if f64 else 1e-6), (k, e)
    assert np.any(ref["R"] != 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32b"])
@pytest.mark.parametrize("name", ["ramp70x30x9", "swept20x12x10", "tgv16"])
def test_fvm_sweep_forms_agree(tmp_path, name, dtype):
    import torch
    import sweep_child
    assert torch.cuda.is_available(), "GPU test"
    case = _case(name, dtype, fvm=f3d.FVM_EIGEN)
    src, dst = _inputs(tmp_path, case, dtype)
    mine = sweep_child.run(f3d, torch, np.load(src))
    print(f"fvm {name} {np.dtype(dtype).name}: child with F3D_FVM_SWEEP_XY=0", flush=True)
    ref = _child(src, dst, "F3D_FVM_SWEEP_XY")
    for k in ("R", "U", "Un", "Q"):
        same = np.array_equal(mine[k], ref[k])
        print(f"  {k}: bit-identical {same}, normwise {relerr(mine[k], ref[k]):.2e}")
        assert same, k
    assert np.any(ref["R"] != 0)


@pytest.mark.parametrize("fvm", [0, 1], ids=["fdm", "fvm"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32b"])
@pytest.mark.parametrize("name", ["ramp70x30x9", "swept20x12x10", "tgv16"])
def test_fill_forms_agree(tmp_path, name, dtype, fvm):
    """The one-launch ghost fill (k_fill_all, the default) against the x, y and z fills in turn
    (F3D_FILL_SPLIT=1): the same values bit for bit, Q and (finite-volume branch) U, after three steps of
    inflow, outflow, isothermal-wall and periodic boundaries (the ramps) and periodic ones (the box)."""
    import torch
    import sweep_child
    assert torch.cuda.is_available(), "GPU test"
    case = _case(name, dtype, fvm=f3d.FVM_EIGEN if fvm else 0)
    src, dst = _inputs(tmp_path, case, dtype)
    mine = sweep_child.run(f3d, torch, np.load(src))
    print(f"{'fvm' if fvm else 'fdm'} {name} {np.dtype(dtype).name}: child with F3D_FILL_SPLIT=1", flush=True)
    ref = _child(src, dst, "F3D_FILL_SPLIT", "1")
    for k in ("R", "U", "Un", "Q"):
        same = np.array_equal(mine[k], ref[k])
        print(f"  {k}: bit-identical {same}")
        assert same, k


@pytest.mark.parametrize("fvm", [0, 1], ids=["fdm", "fvm"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32b"])
@pytest.mark.parametrize("name", ["ramp70x30x9", "tgv16"])
def test_update_schedules_agree(tmp_path, name, dtype, fvm):
    """The viscous-update march's block schedules: plane segments of every tile column (what these small
    grids get by default) against whole columns for the first blocks and segments of the rest
    (F3D_VISC_FULL=3: three whole columns, then segments of the others) and whole columns only
    (F3D_VISC_FULL=1000).  Every plane is marched with the same operands in the same order, so the RHS, U,
    Un and Q agree bit for bit after three steps."""
    import torch
    import sweep_child
    assert torch.cuda.is_available(), "GPU test"
    case = _case(name, dtype, fvm=f3d.FVM_EIGEN if fvm else 0)
    src, dst = _inputs(tmp_path, case, dtype)
    mine = sweep_child.run(f3d, torch, np.load(src))
    for whole in ("3", "1000"):
        print(f"{'fvm' if fvm else 'fdm'} {name} {np.dtype(dtype).name}: child with F3D_VISC_FULL={whole}", flush=True)
        ref = _child(src, dst, "F3D_VISC_FULL", whole)
        for k in ("R", "U", "Un", "Q"):
            same = np.array_equal(mine[k], ref[k])
            print(f"  {k}: bit-identical {same}")
            assert same, (whole, k)
