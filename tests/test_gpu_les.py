"""GPU parity of the LES eddy viscosity (viscous.jl:66-106, params les = Smagorinsky or WALE): the
viscousFlux entry, flowAdvance + div and the fused RHS of both branches, and whole time steps, against
the oracle with the same params.  Bars as in the laminar parity tests (viscous fluxes within 1e-12 of
max|Fv_n| in FP64; FP32-bulk within twice the FP32 oracle's own error against FP64 + 1e-6; the FP64 RHS
within 1e-12 of the long-double oracle or 3x the FP64 oracle's own distance; steps within 10x the
oracle's last-bit sensitivity)."""
import numpy as np
import pytest

from casegen import f3d, fp64_rhs_check, ramp_case, relerr, rhs_scales, scaled_err, tgv_case
from oracle_lib import Oracle
from test_gpu_fvm import fp64_fvm_rhs_check, fvm_rhs_scales, geo_dev
from test_gpu_parity import check_f32b, dev, host, promoted

pytestmark = pytest.mark.gpu

MODELS = [("smag", f3d.LES_SMAG), ("wale", f3d.LES_WALE)]


def _make(name, dtype, **kw):
    if name == "tgv16":
        return tgv_case(16, dtype=dtype, **kw)
    return ramp_case(24, 16, 8, dtype=dtype, jump=1.0, spikes=[(5, 3, 2, 20.0), (15, 8, 5, 3.0)],
                     hybrid_phi3=2.5, **kw)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(params=[np.float64, np.float32], ids=["f64", "f32b"])
def dtype(request):
    return request.param


@pytest.fixture(params=MODELS, ids=[m[0] for m in MODELS])
def les(request):
    return request.param[1]


def _eddy_fraction(case):
    """Share of the viscous-flux cells that the model changes."""
    p = case.params
    lam = type(p).from_buffer_copy(p)
    lam.les = f3d.LES_NONE
    F = [np.zeros(p.fv_shape()) for _ in range(3)]
    Fl = [np.zeros(p.fv_shape()) for _ in range(3)]
    c64 = promoted(case)
    Oracle(p).viscous(*F, c64.Q, c64.metrics, case.J)
    Oracle(lam).viscous(*Fl, c64.Q, c64.metrics, case.J)
    return float(np.mean(np.abs(F[0][0] - Fl[0][0]) > 1e-12 * np.abs(Fl[0][0]).max()))


@pytest.mark.parametrize("cname", ["tgv16", "ramp"])
def test_viscous_flux_les(torch_mod, dtype, les, cname):
    torch = torch_mod
    case = _make(cname, dtype, les=les)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    frac = _eddy_fraction(case)
    assert frac > 0.5, f"the model changes only {frac:.2f} of the cells"
    Fv = [np.zeros(p.fv_shape(), dtype=case.bulk) for _ in range(3)]
    o.viscous(*Fv, case.Q, case.metrics, case.J)
    tdt = torch.float64 if case.bulk == np.float64 else torch.float32
    Fvg = [torch.zeros(p.fv_shape(), dtype=tdt, device="cuda:0") for _ in range(3)]
    ctx.viscousFlux(*Fvg, dev(torch, case.Q), *[dev(torch, m) for m in case.metrics], dev(torch, case.J))
    sc = [max(float(np.max(np.abs(F[n]))) for F in Fv) for n in range(4)]
    if case.bulk == np.float64:
        for d in range(3):
            e = scaled_err(host(Fvg[d]), Fv[d], sc)
            print(f"{cname} les={les} viscous d={d}: {e:.2e} (model active in {frac:.2f} of cells)")
            assert e < 1e-12, (d, e)
        return
    c64 = promoted(case)
    F64 = [np.zeros(p.fv_shape()) for _ in range(3)]
    o.viscous(*F64, c64.Q, c64.metrics, case.J)
    for d in range(3):
        check_f32b(f"{cname} les={les} viscous d={d}", host(Fvg[d]), Fv[d], F64[d], sc)


@pytest.mark.parametrize("cname", ["tgv16", "ramp"])
def test_rhs_les(torch_mod, dtype, les, cname):
    """flowAdvance (kernel-by-kernel) + div and the fused RHS with the eddy viscosity."""
    torch = torch_mod
    case = _make(cname, dtype, les=les)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    Qg = dev(torch, case.Q)
    mg = [dev(torch, m) for m in case.metrics]
    Jg = dev(torch, case.J)
    z = lambda shape: torch.zeros(shape, dtype=Qg.dtype, device="cuda:0")
    F = [z(p.face_shape(d)) for d in range(3)]
    Fv = [z(p.fv_shape()) for _ in range(3)]
    phig = z(p.cell_shape())
    Ug = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.shockSensor(phig, Qg)
    ctx.flowAdvance(Ug, Qg, z(p.cell_shape(5)), z(p.cell_shape(5)), *F, *Fv, *[dev(torch, s) for s in case.s], *mg,
                    Jg, None, None, None, phig)
    ctx.div(Ug, *F, *Fv, 1.0, Jg)
    Rg = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.rhs(Rg, Qg, mg, Jg)
    if case.bulk == np.float64:
        # the model's share of the RHS is far above the bar (it is what the comparison sees)
        lam = type(p).from_buffer_copy(p)
        lam.les = f3d.LES_NONE
        R, Rl = np.zeros(p.cell_shape(5)), np.zeros(p.cell_shape(5))
        o.rhs(R, case.Q, case.metrics, case.J)
        Oracle(lam).rhs(Rl, case.Q, case.metrics, case.J)
        effect = scaled_err(R, Rl, rhs_scales(case))
        assert effect > 1e-7, effect
        fp64_rhs_check(f"{cname} les={les} flowAdvance+div", host(Ug), case)
        fp64_rhs_check(f"{cname} les={les} rhs", host(Rg), case)
        return
    R = np.zeros(p.cell_shape(5))
    o.rhs(R, case.Q, case.metrics, case.J)
    c64 = promoted(case)
    R64 = np.zeros(p.cell_shape(5))
    o.rhs(R64, c64.Q, c64.metrics, c64.J)
    sc = rhs_scales(case)
    check_f32b(f"{cname} les={les} flowAdvance+div", host(Ug), R, R64, sc)
    check_f32b(f"{cname} les={les} rhs", host(Rg), R, R64, sc)


@pytest.mark.parametrize("cname", ["tgv16", "ramp"])
def test_time_step_les(torch_mod, dtype, les, cname):
    torch = torch_mod
    case = _make(cname, dtype, les=les)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    nsteps = 2
    U = case.U()
    Q = case.Q.copy()
    m = case.metrics
    dt = 0.2 * o.cfl_dt(Q, m, case.J, 1.0)
    Uo, Qo = U.copy(), Q.copy()
    o.time_step(nsteps, Uo, np.zeros_like(U), Qo, m, case.J, dt, case.inflow)
    rng = np.random.default_rng(5)
    Up = U * (1 + 1e-16 * (2 * rng.random(U.shape) - 1))
    Qp = Q.copy()
    if case.bulk == np.float32:
        toward = np.where(rng.random(Q.shape) < 0.5, -np.inf, np.inf).astype(np.float32)
        Qp = np.ascontiguousarray(np.nextafter(Q, toward))
        Up = U * (1 + 2.0 ** -24 * (2 * rng.random(U.shape) - 1))
    o.time_step(nsteps, Up, np.zeros_like(U), Qp, m, case.J, dt, case.inflow)
    Ug, Qg = dev(torch, U), dev(torch, Q)
    ctx.time_step(nsteps, Ug, dev(torch, np.zeros_like(U)), Qg, [dev(torch, a) for a in m], dev(torch, case.J), dt,
                  None if case.inflow is None else dev(torch, case.inflow))
    inner = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
    floor = 1e-12 if case.bulk == np.float64 else 1e-6
    sens_u, sens_q = relerr(Up[inner], Uo[inner]), relerr(Qp, Qo)
    e, eq = relerr(host(Ug)[inner], Uo[inner]), relerr(host(Qg), Qo)
    print(f"{cname} les={les} {case.bulk.name}: U err {e:.2e} (sensitivity {sens_u:.2e}), Q err {eq:.2e} "
          f"(sensitivity {sens_q:.2e})")
    assert e <= 10 * sens_u + floor and eq <= 10 * sens_q + floor, (e, eq)


def test_rhs_fvm_les(torch_mod, dtype, les):
    """The finite-volume branch shares the viscous flux: its fused RHS with the eddy viscosity."""
    torch = torch_mod
    case = _make("ramp", dtype, les=les, fvm=f3d.FVM_EIGEN)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    U = case.U()
    Qg = dev(torch, case.Q)
    mg = [dev(torch, m) for m in case.metrics]
    Jg = dev(torch, case.J)
    Rg = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.rhs_fvm(Rg, Qg, dev(torch, U), mg, Jg, geo_dev(torch, ctx, case))
    if case.bulk == np.float64:
        fp64_fvm_rhs_check(f"ramp les={les} rhs_fvm", host(Rg), case, U)
        return
    R = np.zeros(p.cell_shape(5))
    o.rhs_fvm(R, case.Q, U, case.metrics, case.J, case.coords)
    c64 = promoted(case)
    R64 = np.zeros(p.cell_shape(5))
    o.rhs_fvm(R64, c64.Q, U, c64.metrics, c64.J, case.coords)
    check_f32b(f"ramp les={les} rhs_fvm", host(Rg), R, R64, fvm_rhs_scales(c64, U))
