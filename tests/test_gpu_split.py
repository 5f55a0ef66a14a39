"""GPU parity of the flag-selected alternatives of the finite-difference branch: the Lax-Friedrichs,
van Leer and AUSM flux-vector splittings (split.jl:70-255, params split), the component-wise
advect_x/y/z (schemes.jl:86-698, params character = 0), and the fused RHS / multi-step paths run with
every splitting and both reconstructions.

Bars as in the Steger-Warming parity tests: split fluxes within 1e-12 (FP64) / 1e-5 (FP32-bulk)
of max|F_n|; face fluxes within 1e-12 of the face scale (FP64), or within twice the FP32 oracle's
own error against the FP64 oracle plus 1e-6 (FP32-bulk); the FP64 RHS within 1e-12 of the
long-double oracle or 3x the FP64 oracle's own distance from it; steps within 10x the oracle's
response to a last-bit perturbation.
"""
import numpy as np
import pytest

from casegen import f3d, face_scales, fp64_rhs_check, ramp_case, relerr, rhs_scales, scaled_err, tgv_case
from oracle_lib import Oracle
from test_gpu_parity import check_f32b, dev, host, promoted

pytestmark = pytest.mark.gpu

SPLITS = [("lf", f3d.SPLIT_LF), ("vl", f3d.SPLIT_VL), ("ausm", f3d.SPLIT_AUSM)]
RTOL = {np.float64: 1e-12, np.float32: 1e-5}


def _make(name, dtype, **kw):
    if name == "tgv16":
        return tgv_case(16, dtype=dtype, **kw)
    if name == "ramp":
        return ramp_case(24, 16, 8, dtype=dtype, jump=1.0, spikes=[(5, 3, 2, 20.0), (15, 8, 5, 3.0)],
                         hybrid_phi3=2.5, **kw)
    return ramp_case(20, 12, 10, sweep=30.0, dtype=dtype, jump=0.5, **kw)


CASE_NAMES = ["tgv16", "ramp", "swept"]


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(params=[np.float64, np.float32], ids=["f64", "f32b"])
def dtype(request):
    return request.param


@pytest.mark.parametrize("cname", CASE_NAMES)
@pytest.mark.parametrize("sname,method", SPLITS, ids=[s[0] for s in SPLITS])
def test_split_methods(torch_mod, dtype, cname, sname, method):
    torch = torch_mod
    case = _make(cname, dtype)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    tol = RTOL[case.bulk.type]
    Qg = dev(torch, case.Q)
    fn = getattr(ctx, f"fluxSplit_{sname.upper()}")
    for d in range(3):
        A = case.metrics[3 * d: 3 * d + 3]
        Fp = np.zeros(p.cell_shape(5), dtype=case.bulk)
        Fm = np.zeros_like(Fp)
        o.flux_split_method(method, case.Q, Fp, Fm, case.s[d], *A)
        Fpg = torch.full(Fp.shape, 7.0, dtype=Qg.dtype, device="cuda:0")
        Fmg = torch.full(Fm.shape, 7.0, dtype=Qg.dtype, device="cuda:0")
        fn(Qg, Fpg, Fmg, dev(torch, case.s[d]), *[dev(torch, a) for a in A])
        ep, em = relerr(host(Fpg), Fp), relerr(host(Fmg), Fm)
        print(f"{cname} {sname} {case.bulk.name} d={d}: Fp {ep:.2e} Fm {em:.2e}")
        assert ep < tol and em < tol, (d, ep, em)


@pytest.mark.parametrize("cname", CASE_NAMES)
@pytest.mark.parametrize("NV", [5, 3])
def test_component_advect(torch_mod, dtype, cname, NV):
    """advect_x/y/z from the oracle's split fluxes (so the face kernel is checked on its own);
    components NV..4 of F stay as they were."""
    torch = torch_mod
    case = _make(cname, dtype)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    phi = np.zeros(p.cell_shape(), dtype=case.bulk)
    o.shock_sensor(phi, case.Q)
    phig = dev(torch, phi)
    Fo, Fgs, splits = [], [], []
    for d in range(3):
        A = case.metrics[3 * d: 3 * d + 3]
        Fp = np.zeros(p.cell_shape(5), dtype=case.bulk)
        Fm = np.zeros_like(Fp)
        o.flux_split(case.Q, Fp, Fm, case.s[d], *A)
        splits.append((Fp, Fm))
        F = np.full(p.face_shape(d), 3.0, dtype=case.bulk)
        o.advect_comp(d, F, phi, case.s[d], Fp, Fm, NV)
        Fg = torch.full(p.face_shape(d), 3.0, dtype=phig.dtype, device="cuda:0")
        [ctx.advect_x, ctx.advect_y, ctx.advect_z][d](Fg, phig, dev(torch, case.s[d]), dev(torch, Fp),
                                                      dev(torch, Fm), NV)
        Fgs.append(host(Fg))
        Fo.append(F)
        assert np.all(Fgs[d][NV:] == 3.0), "components beyond NV were written"
    sc = face_scales([F[:NV] for F in Fo]) if NV == 5 else [max(float(np.abs(F[n]).max()) for F in Fo)
                                                              for n in range(NV)]
    if case.bulk == np.float64:
        for d in range(3):
            e = scaled_err(Fgs[d][:NV], Fo[d][:NV], sc)
            print(f"{cname} component advect d={d} NV={NV}: face-scaled err {e:.2e}")
            assert e < 1e-12, (d, e)
        return
    c64 = promoted(case)
    phi64 = np.zeros(p.cell_shape())
    o.shock_sensor(phi64, c64.Q)
    for d in range(3):
        Fp, Fm = (np.ascontiguousarray(a.astype(np.float64)) for a in splits[d])
        F64 = np.full(p.face_shape(d), 3.0)
        o.advect_comp(d, F64, phi64, c64.s[d], Fp, Fm, NV)
        check_f32b(f"{cname} component advect d={d} NV={NV}", Fgs[d][:NV], Fo[d][:NV], F64[:NV], sc)


VARIANTS = [(s, c) for s in (f3d.SPLIT_SW, f3d.SPLIT_LF, f3d.SPLIT_VL, f3d.SPLIT_AUSM) for c in (1, 0)
            if not (s == f3d.SPLIT_SW and c == 1)]
VIDS = [f"{['sw', 'lf', 'vl', 'ausm'][s]}_{'char' if c else 'comp'}" for s, c in VARIANTS]


@pytest.mark.parametrize("cname", CASE_NAMES)
@pytest.mark.parametrize("split,character", VARIANTS, ids=VIDS)
def test_flow_advance_and_rhs_variants(torch_mod, dtype, cname, split, character):
    """flowAdvance (kernel-by-kernel) + div and the fused workspace RHS with the selected splitting
    and reconstruction, against the oracle RHS of the same params."""
    torch = torch_mod
    case = _make(cname, dtype, split=split, character=character)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    Qg = dev(torch, case.Q)
    mg = [dev(torch, m) for m in case.metrics]
    Jg = dev(torch, case.J)
    z = lambda shape: torch.zeros(shape, dtype=Qg.dtype, device="cuda:0")
    Fp, Fm = z(p.cell_shape(5)), z(p.cell_shape(5))
    F = [z(p.face_shape(d)) for d in range(3)]
    Fv = [z(p.fv_shape()) for _ in range(3)]
    phig = z(p.cell_shape())
    Ug = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.shockSensor(phig, Qg)
    ctx.flowAdvance(Ug, Qg, Fp, Fm, *F, *Fv, *[dev(torch, s) for s in case.s], *mg, Jg, None, None, None, phig)
    ctx.div(Ug, *F, *Fv, 1.0, Jg)
    Rg = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.rhs(Rg, Qg, mg, Jg)
    if case.bulk == np.float64:
        fp64_rhs_check(f"{cname} {split}/{character} flowAdvance+div", host(Ug), case)
        fp64_rhs_check(f"{cname} {split}/{character} rhs", host(Rg), case)
        return
    R = np.zeros(p.cell_shape(5))
    o.rhs(R, case.Q, case.metrics, case.J)
    c64 = promoted(case)
    R64 = np.zeros(p.cell_shape(5))
    o.rhs(R64, c64.Q, c64.metrics, c64.J)
    sc = rhs_scales(case)
    check_f32b(f"{cname} {split}/{character} flowAdvance+div", host(Ug), R, R64, sc)
    check_f32b(f"{cname} {split}/{character} rhs", host(Rg), R, R64, sc)


@pytest.mark.parametrize("cname", ["tgv16", "ramp"])
@pytest.mark.parametrize("split,character", VARIANTS, ids=VIDS)
def test_time_step_variants(torch_mod, dtype, cname, split, character):
    torch = torch_mod
    case = _make(cname, dtype, split=split, character=character)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    nsteps = 2
    U = case.U()
    Q = case.Q.copy()
    m = case.metrics
    dt = 0.2 * o.cfl_dt(Q, m, case.J, 1.0)
    Uo, Uno, Qo = U.copy(), np.zeros_like(U), Q.copy()
    o.time_step(nsteps, Uo, Uno, Qo, m, case.J, dt, case.inflow)
    rng = np.random.default_rng(5)
    Up = U * (1 + 1e-16 * (2 * rng.random(U.shape) - 1))
    Qp = Q.copy()
    if case.bulk == np.float32:
        toward = np.where(rng.random(Q.shape) < 0.5, -np.inf, np.inf).astype(np.float32)
        Qp = np.ascontiguousarray(np.nextafter(Q, toward))
        Up = U * (1 + 2.0 ** -24 * (2 * rng.random(U.shape) - 1))
    o.time_step(nsteps, Up, np.zeros_like(U), Qp, m, case.J, dt, case.inflow)
    Ug, Ung, Qg = dev(torch, U), dev(torch, np.zeros_like(U)), dev(torch, Q)
    ctx.time_step(nsteps, Ug, Ung, Qg, [dev(torch, a) for a in m], dev(torch, case.J), dt,
                  None if case.inflow is None else dev(torch, case.inflow))
    inner = (slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
    floor = 1e-12 if case.bulk == np.float64 else 1e-6
    sens_u, sens_q = relerr(Up[inner], Uo[inner]), relerr(Qp, Qo)
    e, eq = relerr(host(Ug)[inner], Uo[inner]), relerr(host(Qg), Qo)
    print(f"{cname} {split}/{character} {case.bulk.name}: U err {e:.2e} (sensitivity {sens_u:.2e}), "
          f"Q err {eq:.2e} (sensitivity {sens_q:.2e})")
    assert e <= 10 * sens_u + floor and eq <= 10 * sens_q + floor, (e, eq)
