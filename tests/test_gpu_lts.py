"""GPU parity of local time stepping (params lts): the compute_dt and div_LTS entries (utils.jl:73-100,
div.jl:52-101) against the oracle, and whole time steps of both branches with a per-cell dt.

Bars: compute_dt within 4e-16 relative per cell (the same double operations), ghost cells of dt left
alone; div_LTS increments within 1e-12 (FP64) / 1e-5 (FP32-bulk) of the oracle's, normwise per
variable; steps within 10x the oracle's response to a last-bit perturbation."""
import numpy as np
import pytest

from casegen import f3d, ramp_case, relerr, tgv_case
from oracle_lib import Oracle
from test_gpu_fvm import geo_dev
from test_gpu_parity import dev, host

pytestmark = pytest.mark.gpu
NG = 4


def _make(name, dtype, **kw):
    if name == "tgv16":
        return tgv_case(16, dtype=dtype, lts=1, lts_cfl=0.3, **kw)
    return ramp_case(24, 16, 8, dtype=dtype, jump=1.0, spikes=[(5, 3, 2, 20.0), (15, 8, 5, 3.0)], hybrid_phi3=2.5,
                     lts=1, lts_cfl=0.3, **kw)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(params=[np.float64, np.float32], ids=["f64", "f32b"])
def dtype(request):
    return request.param


@pytest.mark.parametrize("cname", ["tgv16", "ramp"])
def test_compute_dt(torch_mod, dtype, cname):
    torch = torch_mod
    case = _make(cname, dtype)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    Q = case.Q.copy()
    # a few supersonic cells moving towards -x: u + c < 0, negative dt as the reference computes it
    c = np.sqrt(p.gamma * p.Rg * Q[5, 6, 7, 8:12].astype(np.float64))
    Q[1, 6, 7, 8:12] = (-2.0 * c).astype(Q.dtype)
    dto = np.full(p.cell_shape(), -9.0)
    o.compute_dt(dto, Q, case.J, *case.s)
    dtg = torch.full(p.cell_shape(), -9.0, dtype=torch.float64, device="cuda:0")
    ctx.compute_dt(dtg, dev(torch, Q), dev(torch, case.J), *[dev(torch, s) for s in case.s])
    g = host(dtg)
    inner = (slice(NG, -NG),) * 3
    assert np.all(g[6, 7, 8:12] < 0)
    e = np.max(np.abs(g[inner] - dto[inner]) / np.abs(dto[inner]))
    print(f"{cname} {np.dtype(dtype).name} compute_dt: max rel err {e:.2e}")
    assert e <= 4e-16, e
    ghost = np.ones(p.cell_shape(), bool)
    ghost[inner] = False
    assert np.all(g[ghost] == -9.0)


def test_div_lts(torch_mod, dtype):
    torch = torch_mod
    p = f3d.make_params(20, 12, 10)
    o, ctx = Oracle(p), f3d.Flame3D(p)
    rng = np.random.default_rng(8)
    F = [rng.standard_normal(p.face_shape(d)).astype(dtype) for d in range(3)]
    Fv = [rng.standard_normal(p.fv_shape()).astype(dtype) for _ in range(3)]
    J = 1.0 + rng.random(p.cell_shape())
    dtf = 1e-3 * (1 + rng.random(p.cell_shape()))
    U = rng.standard_normal(p.cell_shape(5))
    Uo = np.zeros_like(U)
    o.div_lts(Uo, *F, *Fv, dtf, J)
    Ug = torch.zeros(U.shape, dtype=torch.float64, device="cuda:0")
    ctx.div_LTS(Ug, *[dev(torch, f) for f in F], *[dev(torch, f) for f in Fv], dev(torch, dtf), dev(torch, J))
    e = relerr(host(Ug), Uo)
    print(f"div_LTS {np.dtype(dtype).name}: {e:.2e}")
    assert e < (1e-12 if dtype == np.float64 else 1e-5), e


@pytest.mark.parametrize("cname", ["tgv16", "ramp"])
@pytest.mark.parametrize("branch", ["fdm", "fvm"])
def test_time_step_lts(torch_mod, dtype, cname, branch):
    torch = torch_mod
    kw = dict(fvm=f3d.FVM_EIGEN) if branch == "fvm" else {}
    case = _make(cname, dtype, **kw)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    nsteps = 2
    U = case.U()
    Q = case.Q.copy()
    m = case.metrics
    run = (lambda oo, U_, Q_: oo.time_step_fvm(nsteps, U_, np.zeros_like(U_), Q_, m, case.J, case.coords, 0.0,
                                                 case.inflow)) if branch == "fvm" else \
          (lambda oo, U_, Q_: oo.time_step(nsteps, U_, np.zeros_like(U_), Q_, m, case.J, 0.0, case.inflow))
    Uo, Qo = U.copy(), Q.copy()
    run(o, Uo, Qo)
    rng = np.random.default_rng(5)
    Up = U * (1 + 1e-16 * (2 * rng.random(U.shape) - 1))
    Qp = Q.copy()
    if case.bulk == np.float32:
        toward = np.where(rng.random(Q.shape) < 0.5, -np.inf, np.inf).astype(np.float32)
        Qp = np.ascontiguousarray(np.nextafter(Q, toward))
        Up = U * (1 + 2.0 ** -24 * (2 * rng.random(U.shape) - 1))
    run(o, Up, Qp)
    Ug, Qg = dev(torch, U), dev(torch, Q)
    mg = [dev(torch, a) for a in m]
    inflow = None if case.inflow is None else dev(torch, case.inflow)
    # the scalar dt is ignored under lts: pass a nonsense value
    if branch == "fvm":
        ctx.time_step_fvm(nsteps, Ug, dev(torch, np.zeros_like(U)), Qg, mg, dev(torch, case.J),
                          geo_dev(torch, ctx, case), 1e9, inflow)
    else:
        ctx.time_step(nsteps, Ug, dev(torch, np.zeros_like(U)), Qg, mg, dev(torch, case.J), 1e9, inflow)
    inner = (slice(None),) + (slice(NG, -NG),) * 3
    floor = 1e-12 if case.bulk == np.float64 else 1e-6
    sens_u, sens_q = relerr(Up[inner], Uo[inner]), relerr(Qp, Qo)
    e, eq = relerr(host(Ug)[inner], Uo[inner]), relerr(host(Qg), Qo)
    assert relerr(Uo[inner], U[inner]) > 1e-6, "the steps did not move the state"
    print(f"{cname} {branch} lts {case.bulk.name}: U err {e:.2e} (sensitivity {sens_u:.2e}), Q err {eq:.2e} "
          f"(sensitivity {sens_q:.2e})")
    assert e <= 10 * sens_u + floor and eq <= 10 * sens_q + floor, (e, eq)
