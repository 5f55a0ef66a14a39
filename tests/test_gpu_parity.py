"""GPU parity: every C-ABI entry point against the CPU oracle on identical seeded inputs.

Bar: the integer scheme mask and the sensor bit-exact; copies (ghost fill, exchange) bit-exact;
floating-point results within RTOL of the magnitude of the terms they are summed from:
  face fluxes  max|gpu - oracle|_n / max_{faces, 3 directions}|F_n|           (face_scales)
  RHS          max|gpu - oracle|_n / max_cells J*sum|flux and viscous terms|_n  (rhs_scales)
(the characteristic projection mixes rho*c^2-sized terms into fluxes that can vanish, e.g. the
zeta-face mass flux of a sheared grid, so max|F_n| of one direction alone is no scale).
RTOL = 1e-12 for the FP64 face fluxes.  The FP64 RHS is compared with a long-double evaluation of
the same formulas (mask frozen): within 1e-12 of it, or no worse than 3x the FP64 oracle's own
distance from it where the input's FP64 floor approaches 1e-12 (a pressure spike at the wall).
The FP32-bulk form is judged against the FP64 oracle run on the same (float-rounded) inputs: its
error may not exceed twice the FP32-bulk oracle's own error against that FP64 result, plus 1e-6
(i.e. the GPU FP32-bulk path is as accurate as the FP32 restatement of the reference).
Multi-step results are compared with a bar set by the oracle's own response, after the same
steps, to a 1e-16 relative perturbation of U (10x that response + 1e-12; FP32-bulk: a one-ulp
perturbation of Q, 10x + 1e-6): after the first stage the shock sensor and WENO weights see
last-bit differences, and the scheme amplifies them far beyond rounding.
"""
import numpy as np
import pytest

from casegen import f3d, face_scales, fp64_rhs_check, ramp_case, relerr, rhs_scales, scaled_err, tgv_case
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

RTOL = {np.float64: 1e-12, np.float32: 1e-5}


def _cases():
    return [
        ("tgv16", lambda dt: tgv_case(16, dtype=dt)),
        ("tgv16_nnd", lambda dt: tgv_case(16, dtype=dt, hybrid_phi3=2.5)),
        ("ramp", lambda dt: ramp_case(24, 16, 8, dtype=dt, jump=1.0, spikes=[(5, 3, 2, 20.0), (15, 8, 5, 3.0)],
                                      hybrid_phi3=2.5)),
        ("swept", lambda dt: ramp_case(20, 12, 10, sweep=30.0, dtype=dt, jump=0.5)),
    ]


def promoted(case):
    """The case's float inputs promoted to float64 (the FP64 reference for FP32-bulk checks)."""
    import copy
    c = copy.copy(case)
    c.Q = np.ascontiguousarray(case.Q.astype(np.float64))
    c.metrics = [np.ascontiguousarray(m.astype(np.float64)) for m in case.metrics]
    c.s = [np.ascontiguousarray(a.astype(np.float64)) for a in case.s]
    c.inflow = None if case.inflow is None else np.ascontiguousarray(case.inflow.astype(np.float64))
    return c


def check_f32b(name, gpu, o32, o64, scale):
    e_floor = scaled_err(o32, o64, scale)
    e_gpu = scaled_err(gpu, o64, scale)
    print(f"{name}: FP32-bulk gpu-vs-fp64 {e_gpu:.2e}, oracle32-vs-fp64 {e_floor:.2e}")
    assert e_gpu <= 2 * e_floor + 1e-6, f"{name}: {e_gpu:.3e} > 2*{e_floor:.3e}+1e-6"


CASES = _cases()
IDS = [c[0] for c in CASES]


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(params=[np.float64, np.float32], ids=["f64", "f32b"])
def dtype(request):
    return request.param


@pytest.fixture(params=CASES, ids=IDS)
def case(request, dtype):
    return request.param[1](dtype)


def test_sensor_and_mask_bit_exact(torch_mod, case):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    phi_o = np.zeros(p.cell_shape(), dtype=case.bulk)
    o.shock_sensor(phi_o, case.Q)
    phi_g = torch.zeros(p.cell_shape(), dtype=torch.float64 if case.bulk == np.float64 else torch.float32,
                        device="cuda:0")
    ctx.shockSensor(phi_g, dev(torch, case.Q))
    assert np.array_equal(host(phi_g), phi_o), "shock sensor differs bit-wise"
    mo = [np.zeros(p.mask_shape(d), np.uint8) for d in range(3)]
    o.scheme_mask(*mo, phi_o)
    mg = [torch.full(p.mask_shape(d), 255, dtype=torch.uint8, device="cuda:0") for d in range(3)]
    ctx.schemeMask(*mg, phi_g)
    for d in range(3):
        assert np.array_equal(host(mg[d]), mo[d]), f"scheme mask differs in direction {d}"


def test_split_and_advect(torch_mod, case):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    tol = RTOL[case.bulk.type]
    phi = np.zeros(p.cell_shape(), dtype=case.bulk)
    o.shock_sensor(phi, case.Q)
    Qg, phig = dev(torch, case.Q), dev(torch, phi)
    Fo, Fgs = [], []
    for d in range(3):
        A = case.metrics[3 * d: 3 * d + 3]
        S = case.s[d]
        Fp = np.zeros(p.cell_shape(5), dtype=case.bulk)
        Fm = np.zeros_like(Fp)
        o.flux_split(case.Q, Fp, Fm, S, *A)
        Fpg, Fmg = torch.zeros_like(dev(torch, Fp)), torch.zeros_like(dev(torch, Fm))
        Ag = [dev(torch, a) for a in A]
        Sg = dev(torch, S)
        ctx.fluxSplit_SW(Qg, Fpg, Fmg, Sg, *Ag)
        assert relerr(host(Fpg), Fp) < tol and relerr(host(Fmg), Fm) < tol, f"split d={d}"
        F = np.zeros(p.face_shape(d), dtype=case.bulk)
        o.advect(d, F, phi, S, Fp, Fm, case.Q, *A)
        # advect from the oracle's split fluxes so the face kernel is checked on its own
        Fg = torch.zeros(p.face_shape(d), dtype=Qg.dtype, device="cuda:0")
        [ctx.advect_xc, ctx.advect_yc, ctx.advect_zc][d](Fg, phig, Sg, dev(torch, Fp), dev(torch, Fm), Qg, *Ag)
        Fo.append(F)
        Fgs.append(host(Fg))
    sc = face_scales(Fo)
    if case.bulk == np.float64:
        for d in range(3):
            e = scaled_err(Fgs[d], Fo[d], sc)
            print(f"{case.name} {case.bulk} advect d={d}: face-scaled err {e:.2e}")
            assert e < tol, f"advect d={d}: {e}"
        return
    c64 = promoted(case)
    phi64 = np.zeros(p.cell_shape())
    o.shock_sensor(phi64, c64.Q)
    for d in range(3):
        A = c64.metrics[3 * d: 3 * d + 3]
        Fp = np.zeros(p.cell_shape(5))
        Fm = np.zeros_like(Fp)
        o.flux_split(c64.Q, Fp, Fm, c64.s[d], *A)
        F64 = np.zeros(p.face_shape(d))
        o.advect(d, F64, phi64, c64.s[d], Fp, Fm, c64.Q, *A)
        check_f32b(f"{case.name} advect d={d}", Fgs[d], Fo[d], F64, sc)


def test_viscous_and_div(torch_mod, case):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    tol = RTOL[case.bulk.type]
    Fv = [np.zeros(p.fv_shape(), dtype=case.bulk) for _ in range(3)]
    o.viscous(*Fv, case.Q, case.metrics, case.J)
    Fvg = [torch.zeros(p.fv_shape(), dtype=torch.float64 if case.bulk == np.float64 else torch.float32,
                       device="cuda:0") for _ in range(3)]
    mg = [dev(torch, m) for m in case.metrics]
    Jg = dev(torch, case.J)
    ctx.viscousFlux(*Fvg, dev(torch, case.Q), *mg, Jg)
    fvs = [max(float(np.max(np.abs(F[n]))) for F in Fv) for n in range(4)]
    for d in range(3):
        e = scaled_err(host(Fvg[d]), Fv[d], fvs)
        assert e < tol, f"viscous d={d}: {e}"
    rng = np.random.default_rng(7)
    F = [rng.standard_normal(p.face_shape(d)).astype(case.bulk) for d in range(3)]
    U = rng.standard_normal(p.cell_shape(5))
    Uo = U.copy()
    o.div(Uo, *F, *Fv, 1e-3, case.J)
    Ug = dev(torch, U)
    ctx.div(Ug, *[dev(torch, f) for f in F], *[dev(torch, f) for f in Fv], 1e-3, Jg)
    e = relerr(host(Ug) - U, Uo - U)
    assert e < tol, f"div increment: {e}"


def test_conversions_and_lincomb(torch_mod, case):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    U = case.U()
    Qo = np.zeros_like(case.Q)
    o.c2prim(U, Qo)
    Qg = torch.zeros_like(dev(torch, case.Q))
    ctx.c2Prim(dev(torch, U), Qg)
    tol = 1e-14 if case.bulk == np.float64 else 1e-7
    assert relerr(host(Qg), Qo) < tol
    Uo = np.zeros_like(U)
    o.prim2c(Uo, case.Q)
    Ug = torch.zeros_like(dev(torch, U))
    ctx.prim2c(Ug, dev(torch, case.Q))
    assert relerr(host(Ug), Uo) < 1e-14
    Un = np.random.default_rng(3).standard_normal(U.shape)
    Ul = U.copy()
    o.lin_comb(Ul, Un, 5, 0.25, 0.75)
    Ulg = dev(torch, U)
    ctx.linComb(Ulg, dev(torch, Un), 5, 0.25, 0.75)
    assert relerr(host(Ulg), Ul) < 1e-15


def test_ghost_fill_bit_exact(torch_mod, case):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    rng = np.random.default_rng(11)
    Q = (case.Q * (1 + 0.01 * rng.random(case.Q.shape))).astype(case.bulk)
    U = rng.random(p.cell_shape(5))
    Qo, Uo = Q.copy(), U.copy()
    o.exchange(Qo, 6)
    o.fill(Qo, Uo, case.inflow)
    Qg, Ug = dev(torch, Q), dev(torch, U)
    ctx.exchange_ghost(Qg, 6)
    ctx.fillGhost(Qg, Ug, None if case.inflow is None else dev(torch, case.inflow))
    assert np.array_equal(host(Qg), Qo), "ghost fill of Q differs"
    # U ghosts computed from Q in double: equal up to rounding of the same expression
    assert relerr(host(Ug), Uo) < 1e-15


def test_flow_advance_and_rhs(torch_mod, case):
    """flowAdvance (kernel-by-kernel) + div, and the workspace RHS, against the oracle RHS."""
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    tol = RTOL[case.bulk.type]
    R = np.zeros(p.cell_shape(5))
    phi = np.zeros(p.cell_shape(), dtype=case.bulk)
    o.rhs(R, case.Q, case.metrics, case.J, phi)
    Qg = dev(torch, case.Q)
    mg = [dev(torch, m) for m in case.metrics]
    Jg = dev(torch, case.J)
    tdt = Qg.dtype
    z = lambda shape: torch.zeros(shape, dtype=tdt, device="cuda:0")
    Fp, Fm = z(p.cell_shape(5)), z(p.cell_shape(5))
    F = [z(p.face_shape(d)) for d in range(3)]
    Fv = [z(p.fv_shape()) for _ in range(3)]
    phig = z(p.cell_shape())
    sg = [dev(torch, s) for s in case.s]
    Ug = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.shockSensor(phig, Qg)
    ctx.flowAdvance(Ug, Qg, Fp, Fm, *F, *Fv, *sg, *mg, Jg, None, None, None, phig)
    ctx.div(Ug, *F, *Fv, 1.0, Jg)
    sc = rhs_scales(case)
    Rg = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.rhs(Rg, Qg, mg, Jg)
    if case.bulk == np.float64:
        fp64_rhs_check(f"{case.name} flowAdvance+div", host(Ug), case)
        fp64_rhs_check(f"{case.name} rhs", host(Rg), case)
        print(f"{case.name} plain normwise rhs err {relerr(host(Rg), R):.2e}")
        return
    c64 = promoted(case)
    R64 = np.zeros(p.cell_shape(5))
    o.rhs(R64, c64.Q, c64.metrics, c64.J)
    check_f32b(f"{case.name} flowAdvance+div", host(Ug), R, R64, sc)
    check_f32b(f"{case.name} rhs", host(Rg), R, R64, sc)


@pytest.mark.parametrize("nsteps", [1, 2, 3])
def test_time_step(torch_mod, case, nsteps):
    """nsteps RK3 steps in one call against the oracle's time_step: U and Q, and Un, which holds the last
    step's U^n on return as in the reference's loop (the library stages the other steps through Un)."""
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    U = case.U()
    Un = np.zeros_like(U)
    Q = case.Q.copy()
    m = case.metrics
    dt = 0.2 * o.cfl_dt(Q, m, case.J, 1.0)
    Uo, Uno, Qo = U.copy(), Un.copy(), Q.copy()
    o.time_step(nsteps, Uo, Uno, Qo, m, case.J, dt, case.inflow)
    # the oracle's own response to a last-bit perturbation of the initial state
    rng = np.random.default_rng(5)
    if case.bulk == np.float64:
        Up = U * (1 + 1e-16 * (2 * rng.random(U.shape) - 1))
        Qp = Q.copy()
    else:
        Up = U.copy()
        toward = np.where(rng.random(Q.shape) < 0.5, -np.inf, np.inf).astype(np.float32)
        Qp = np.ascontiguousarray(np.nextafter(Q, toward))
    Unp = np.zeros_like(U)
    o.time_step(nsteps, Up, Unp, Qp, m, case.J, dt, case.inflow)
    Ug, Ung, Qg = dev(torch, U), dev(torch, Un), dev(torch, Q)
    ctx.time_step(nsteps, Ug, Ung, Qg, [dev(torch, a) for a in m], dev(torch, case.J), dt,
                  None if case.inflow is None else dev(torch, case.inflow))
    NG = 4
    inner = (slice(None), slice(NG, -NG), slice(NG, -NG), slice(NG, -NG))
    floor = 1e-12 if case.bulk == np.float64 else 1e-6
    sens_u = relerr(Up[inner], Uo[inner])
    sens_q = relerr(Qp, Qo)
    e = relerr(host(Ug)[inner], Uo[inner])
    eq = relerr(host(Qg), Qo)
    eun = relerr(host(Ung)[inner], Uno[inner])
    sens_un = relerr(Unp[inner], Uno[inner])
    print(f"{case.name} {case.bulk} {nsteps} steps: U err {e:.2e} (oracle sensitivity {sens_u:.2e}), "
          f"Q err {eq:.2e} (sensitivity {sens_q:.2e}), Un err {eun:.2e} (sensitivity {sens_un:.2e})")
    assert e <= 10 * sens_u + floor, f"U after {nsteps} steps: {e}"
    assert eq <= 10 * sens_q + floor, f"Q after {nsteps} steps: {eq}"
    assert eun <= 10 * sens_un + floor, f"Un after {nsteps} steps: {eun}"


def test_cfl_and_nan(torch_mod, case):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    ref = o.cfl_dt(case.Q, case.metrics, case.J, 0.5)
    out = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    ctx.cfl_dt(out, dev(torch, case.Q), [dev(torch, a) for a in case.metrics], dev(torch, case.J), 0.5)
    got = float(host(out)[0])
    assert abs(got - ref) <= 1e-12 * ref
    Ug = dev(torch, case.U())
    ctx.nan_check(Ug)
    Ug[2, 5, 6, 7] = float("nan")
    with pytest.raises(f3d.F3DNaNError):
        ctx.nan_check(Ug)


def test_metrics_fp64(torch_mod, cases):
    torch = torch_mod
    c = cases.ramp(20, 12, 10, sweep=30.0)
    p = f3d.make_params(20, 12, 10, **c["params"])
    o = Oracle(p)
    m = [np.zeros(p.cell_shape()) for _ in range(9)]
    J = np.zeros(p.cell_shape())
    o.metrics(c["x"], c["y"], c["z"], m, J)
    mg = [torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0") for _ in range(9)]
    Jg = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    f3d.Flame3D(p).metrics(dev(torch, c["x"]), dev(torch, c["y"]), dev(torch, c["z"]), mg, Jg)
    for q in range(9):
        assert relerr(host(mg[q]), m[q]) < 1e-12, f"metric {q}"
    assert relerr(host(Jg), J) < 1e-12


@pytest.mark.parametrize("bc", [[5, 2, 4, 3, 3, 4], [1, 2, 3, 1, 4, 2], [2, 1, 1, 4, 3, 3]],
                         ids=["inflow_walls", "dirichlet_mix", "zerograd_walls"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32b"])
def test_ghost_fill_every_kind(torch_mod, bc, dtype):
    """The one-launch ghost fill against the oracle's x, y, z fills in turn, on boundary sets that put every
    kind (inflow, Dirichlet, zero gradient, adiabatic and isothermal walls) on faces of all three directions,
    so that the edge and corner ghosts compose different kinds: Q bit for bit, U (from Q in double) to
    rounding."""
    torch = torch_mod
    case = ramp_case(20, 12, 10, dtype=dtype, jump=0.5, bc=bc)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    rng = np.random.default_rng(13)
    Q = (case.Q * (1 + 0.01 * rng.random(case.Q.shape))).astype(case.bulk)
    U = rng.random(p.cell_shape(5))
    Qo, Uo = Q.copy(), U.copy()
    o.fill(Qo, Uo, case.inflow)
    Qg, Ug = dev(torch, Q), dev(torch, U)
    ctx.fillGhost(Qg, Ug, None if case.inflow is None else dev(torch, case.inflow))
    assert np.array_equal(host(Qg), Qo), "ghost fill of Q differs"
    assert relerr(host(Ug), Uo) < 1e-15
