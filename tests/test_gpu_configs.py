"""The CPU-runnable configuration at full size: Taylor-Green box 64^3 (identity metrics, all
periodic, M 0.1, Re 1600), with the parity perturbations (seeded 1e-3 noise, a planted pressure
jump so that every scheme branch fires): one RHS evaluation against the long-double oracle, 100
FP64 RK3 steps at CFL 0.5 against the oracle's 100 steps, and the FP32-bulk form's density drift.

Bars as in test_gpu_parity: the RHS within 1e-12 (term-scaled) of the long-double evaluation (or 3x
the FP64 oracle's own distance); 100 steps within 10x the oracle's own response to a 1e-16
relative perturbation of U, plus 1e-12; the FP32-bulk density drift from FP64 (L2, relative)
no more than twice the FP32-bulk oracle's own drift over the same steps, plus 1e-6 (the spikes
and the jump make this case far rougher than the ramp, where the drift is ~1e-7)."""
import numpy as np
import pytest

from casegen import f3d, fp64_rhs_check, relerr, tgv_case
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

N = 64


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def tgv64():
    spikes = [(9, 17, 33, 20.0), (40, 12, 50, 3.0)]
    return tgv_case(N, noise=1e-3, jump=0.6, seed=1, spikes=spikes)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_tgv64_rhs(torch_mod, tgv64):
    torch = torch_mod
    c = tgv64
    p = c.params
    ctx = f3d.Flame3D(p)
    R = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.rhs(R, dev(torch, c.Q), [dev(torch, m) for m in c.metrics], dev(torch, c.J))
    torch.cuda.synchronize()
    fp64_rhs_check("tgv64", R.cpu().numpy(), c)


def test_tgv64_100_steps(torch_mod, tgv64):
    torch = torch_mod
    c = tgv64
    p, o = c.params, Oracle(c.params)
    U0 = c.U()
    dt = o.cfl_dt(c.Q, c.metrics, c.J, 0.5)
    nsteps = 100
    Uo, Qo = U0.copy(), c.Q.copy()
    o.time_step(nsteps, Uo, np.zeros_like(U0), Qo, c.metrics, c.J, dt)
    rng = np.random.default_rng(5)
    Up = U0 * (1 + 1e-16 * (2 * rng.random(U0.shape) - 1))
    Qp = c.Q.copy()
    o.time_step(nsteps, Up, np.zeros_like(U0), Qp, c.metrics, c.J, dt)
    U32o, Q32o = U0.copy(), np.ascontiguousarray(c.Q.astype(np.float32))
    m32 = [np.ascontiguousarray(a.astype(np.float32)) for a in c.metrics]
    Oracle(p).time_step(nsteps, U32o, np.zeros_like(U0), Q32o, m32, c.J, dt)
    m = [dev(torch, a) for a in c.metrics]
    J = dev(torch, c.J)
    ctx = f3d.Flame3D(p)
    Ug, Qg = dev(torch, U0), dev(torch, c.Q)
    ctx.time_step(nsteps, Ug, torch.zeros_like(Ug), Qg, m, J, dt)
    # FP32-bulk from the same state, same dt
    U32, Q32 = dev(torch, U0), dev(torch, c.Q.astype(np.float32))
    ctx.time_step(nsteps, U32, torch.zeros_like(U32), Q32, [a.float() for a in m], J, dt)
    torch.cuda.synchronize()
    inner = (slice(None),) + (slice(4, -4),) * 3
    e = relerr(Ug.cpu().numpy()[inner], Uo[inner])
    sens = relerr(Up[inner], Uo[inner])
    r64, r32 = Ug[0][inner[1:]], U32[0][inner[1:]]
    drift = float(torch.linalg.norm(r32 - r64) / torch.linalg.norm(r64))
    ro64, ro32 = Uo[0][inner[1:]], U32o[0][inner[1:]]
    drift_o = float(np.linalg.norm(ro32 - ro64) / np.linalg.norm(ro64))
    print(f"tgv64 {nsteps} steps: U err {e:.2e} (oracle sensitivity {sens:.2e}); FP32-bulk rho L2 {drift:.2e} "
          f"(FP32-bulk oracle {drift_o:.2e})")
    assert e <= 10 * sens + 1e-12
    assert drift <= 2 * drift_o + 1e-6
