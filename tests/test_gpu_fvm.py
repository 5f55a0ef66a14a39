"""GPU parity of the finite-volume branch (FVM.jl) against the CPU oracle on identical seeded inputs.

Bars as in test_gpu_parity.py:
  face fluxes (Eigen_/Conser_reconstruct_{i,j,k})  FP64: max|gpu - F_ld|_n / max over the faces of
      the three directions |F_n| within 1e-12, or no worse than 3x the FP64 oracle's own distance from
      F_ld, a long-double evaluation of the same formulas (WENO weights of stencils that are flat to
      a few ulps are set by rounding noise); FP32-bulk: no more than twice the FP32 oracle's own
      error against the FP64 oracle (promoted inputs) + 1e-6.
  RHS (flowAdvance + div, and the fused f3d_rhs_fvm)  FP64: within 1e-12 of a long-double evaluation
      (mask frozen), term-scaled by J * sum of |face fluxes| and viscous terms per cell, or no worse
      than 3x the FP64 oracle's own distance from it; FP32-bulk as above.
  time steps  U (interior and ghosts) and Q within 10x the oracle's own response to a last-bit
      perturbation of the initial state (FP64: U by 1e-16 relative, +1e-12; FP32-bulk: Q by one ulp
      and U by 2^-24 relative, as U is reconstructed in float, +1e-6).
"""
import numpy as np
import pytest

from casegen import f3d, face_scales, ramp_case, relerr, scaled_err, tgv_case
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

NG = 4


def _cases():
    return [
        ("tgv16", lambda dt, fvm: tgv_case(16, dtype=dt, fvm=fvm)),
        ("ramp", lambda dt, fvm: ramp_case(24, 16, 8, dtype=dt, jump=1.0, spikes=[(5, 3, 2, 20.0), (15, 8, 5, 3.0)],
                                           hybrid_phi3=2.5, fvm=fvm)),
        ("swept", lambda dt, fvm: ramp_case(20, 12, 10, sweep=30.0, dtype=dt, jump=0.5, fvm=fvm)),
    ]


CASES = _cases()


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


@pytest.fixture(params=[f3d.FVM_EIGEN, f3d.FVM_CONSER], ids=["eigen", "conser"])
def fvm(request):
    return request.param


@pytest.fixture(params=CASES, ids=[c[0] for c in CASES])
def case(request, dtype, fvm):
    return request.param[1](dtype, fvm)


def promoted(case):
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


def geo_dev(torch, ctx, case):
    """Face geometry arrays of the case's mesh, computed on the GPU (f3d_fvm_geometry_*)."""
    p = case.params
    tdt = torch.float64 if case.bulk == np.float64 else torch.float32
    geo = [torch.zeros((4,) + p.face_shape(d)[1:], dtype=tdt, device="cuda:0") for d in range(3)]
    ctx.fvm_geometry(*[dev(torch, a) for a in case.coords], geo)
    return geo


def fvm_rhs_scales(case, U):
    A = np.zeros(case.params.cell_shape(5))
    Oracle(case.params).rhs_absterms_fvm(A, case.Q, U, case.metrics, case.J, case.coords)
    return [float(np.max(A[n])) for n in range(5)]


def fp64_fvm_rhs_check(name, gpu_R, case, U):
    p, o = case.params, Oracle(case.params)
    R = np.zeros(p.cell_shape(5))
    Rl = np.zeros_like(R)
    o.rhs_fvm(R, case.Q, U, case.metrics, case.J, case.coords)
    o.rhs_fvm_ld(Rl, case.Q, U, case.metrics, case.J, case.coords)
    sc = fvm_rhs_scales(case, U)
    e_gpu, e_or = scaled_err(gpu_R, Rl, sc), scaled_err(R, Rl, sc)
    print(f"{name}: FP64 FVM RHS vs long double {e_gpu:.2e} (oracle {e_or:.2e}; gpu vs oracle "
          f"{scaled_err(gpu_R, R, sc):.2e})")
    assert e_gpu <= max(1e-12, 3 * e_or), f"{name}: {e_gpu:.3e} vs oracle floor {e_or:.3e}"


def _scheme_counts(case):
    p, o = case.params, Oracle(case.params)
    phi = np.zeros(p.cell_shape(), dtype=case.bulk)
    o.shock_sensor(phi, case.Q)
    mk = [np.zeros(p.mask_shape(d), np.uint8) for d in range(3)]
    o.scheme_mask(*mk, phi)
    return phi, np.bincount(np.concatenate([m.ravel() for m in mk]), minlength=4)


def test_reconstruct_faces(torch_mod, case):
    """Eigen_/Conser_reconstruct_{i,j,k} one by one against the oracle."""
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    eigen = p.fvm == f3d.FVM_EIGEN
    U = case.U()
    phi, counts = _scheme_counts(case)
    print(f"{case.name} faces per scheme (UP7, WENO7, WENO5, NND2): {counts.tolist()}")
    Qg, Ug, phig = dev(torch, case.Q), dev(torch, U), dev(torch, phi)
    Xg = [dev(torch, a) for a in case.coords]
    name = "Eigen" if eigen else "Conser"
    Fo, Fgs = [], []
    for d in range(3):
        F = np.zeros(p.face_shape(d), dtype=case.bulk)
        o.fvm_reconstruct(d, eigen, case.Q, U, phi, case.s[d], F, *case.coords)
        Fg = torch.full(p.face_shape(d), float("nan"), dtype=Qg.dtype, device="cuda:0")
        getattr(ctx, f"{name}_reconstruct_{'ijk'[d]}")(Qg, Ug, phig, dev(torch, case.s[d]), Fg, *Xg)
        Fo.append(F)
        Fgs.append(host(Fg))
    sc = face_scales(Fo)
    if case.bulk == np.float64:
        for d in range(3):
            Fl = np.zeros(p.face_shape(d))
            o.fvm_reconstruct_ld(d, eigen, case.Q, U, phi, case.s[d], Fl, *case.coords)
            e, e_or = scaled_err(Fgs[d], Fl, sc), scaled_err(Fo[d], Fl, sc)
            print(f"{case.name} {name} d={d}: face-scaled err vs long double {e:.2e} (oracle {e_or:.2e}, "
                  f"gpu vs oracle {scaled_err(Fgs[d], Fo[d], sc):.2e})")
            assert e <= max(1e-12, 3 * e_or), f"{name} d={d}: {e:.3e} vs oracle floor {e_or:.3e}"
        return
    c64 = promoted(case)
    phi64 = np.zeros(p.cell_shape())
    o.shock_sensor(phi64, c64.Q)
    for d in range(3):
        F64 = np.zeros(p.face_shape(d))
        o.fvm_reconstruct(d, eigen, c64.Q, U, phi64, c64.s[d], F64, *case.coords)
        check_f32b(f"{case.name} {name} d={d}", Fgs[d], Fo[d], F64, sc)


def test_face_geometry(torch_mod, case):
    """f3d_fvm_geometry_* against the oracle's face geometry (FVM.jl:16-31 order of summation there,
    differences first on the GPU; on the ramp's stretched mesh the reference order loses ~1e-13 to
    the cancellation of nearly equal coordinates): FP64 within 1e-12 (normals, and areas relative to
    the largest); FP32-bulk: the FP64 values rounded to float."""
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    geo = geo_dev(torch, ctx, case)
    for d in range(3):
        g = np.zeros((4,) + p.face_shape(d)[1:])
        o.fvm_geometry(d, *case.coords, g)
        h = host(geo[d]).astype(np.float64)
        tol = 1e-12 if case.bulk == np.float64 else 1e-7
        assert np.abs(h[:3] - g[:3]).max() < tol, f"normals d={d}"
        assert np.abs(h[3] - g[3]).max() < tol * np.abs(g[3]).max(), f"areas d={d}"


def test_flow_advance_and_rhs_fvm(torch_mod, case):
    """flowAdvance's finite-volume branch + div, and the fused f3d_rhs_fvm, against the oracle RHS."""
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    U = case.U()
    Qg, Ug0 = dev(torch, case.Q), dev(torch, U)
    mg = [dev(torch, m) for m in case.metrics]
    Jg = dev(torch, case.J)
    Xg = [dev(torch, a) for a in case.coords]
    tdt = Qg.dtype
    z = lambda shape: torch.zeros(shape, dtype=tdt, device="cuda:0")
    F = [z(p.face_shape(d)) for d in range(3)]
    Fv = [z(p.fv_shape()) for _ in range(3)]
    phig = z(p.cell_shape())
    sg = [dev(torch, s) for s in case.s]
    ctx.shockSensor(phig, Qg)
    ctx.flowAdvance(Ug0, Qg, None, None, *F, *Fv, *sg, *mg, Jg, *Xg, phig)
    Rk = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.div(Rk, *F, *Fv, 1.0, Jg)
    Rg = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    ctx.rhs_fvm(Rg, Qg, Ug0, mg, Jg, geo_dev(torch, ctx, case))
    assert np.array_equal(host(Ug0), U), "rhs_fvm / flowAdvance modified U"
    if case.bulk == np.float64:
        fp64_fvm_rhs_check(f"{case.name} fvm={p.fvm} flowAdvance+div", host(Rk), case, U)
        fp64_fvm_rhs_check(f"{case.name} fvm={p.fvm} rhs_fvm", host(Rg), case, U)
        return
    R = np.zeros(p.cell_shape(5))
    o.rhs_fvm(R, case.Q, U, case.metrics, case.J, case.coords)
    c64 = promoted(case)
    R64 = np.zeros(p.cell_shape(5))
    o.rhs_fvm(R64, c64.Q, U, c64.metrics, c64.J, case.coords)
    sc = fvm_rhs_scales(c64, U)
    check_f32b(f"{case.name} fvm={p.fvm} flowAdvance+div", host(Rk), R, R64, sc)
    check_f32b(f"{case.name} fvm={p.fvm} rhs_fvm", host(Rg), R, R64, sc)


@pytest.mark.parametrize("nsteps", [1, 2])
def test_time_step_fvm(torch_mod, case, nsteps):
    torch = torch_mod
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    U = case.U()
    Un = np.zeros_like(U)
    Q = case.Q.copy()
    m = case.metrics
    dt = 0.2 * o.cfl_dt(Q, m, case.J, 1.0)
    Uo, Uno, Qo = U.copy(), Un.copy(), Q.copy()
    o.time_step_fvm(nsteps, Uo, Uno, Qo, m, case.J, case.coords, dt, case.inflow)
    rng = np.random.default_rng(5)
    if case.bulk == np.float64:
        Up = U * (1 + 1e-16 * (2 * rng.random(U.shape) - 1))
        Qp = Q.copy()
    else:
        # the finite-volume branch reconstructs U in float: perturb it at float resolution too
        Up = U * (1 + 2.0 ** -24 * (2 * rng.random(U.shape) - 1))
        toward = np.where(rng.random(Q.shape) < 0.5, -np.inf, np.inf).astype(np.float32)
        Qp = np.ascontiguousarray(np.nextafter(Q, toward))
    Unp = np.zeros_like(U)
    o.time_step_fvm(nsteps, Up, Unp, Qp, m, case.J, case.coords, dt, case.inflow)
    Ug, Ung, Qg = dev(torch, U), dev(torch, Un), dev(torch, Q)
    ctx.time_step_fvm(nsteps, Ug, Ung, Qg, [dev(torch, a) for a in m], dev(torch, case.J), geo_dev(torch, ctx, case),
                      dt, None if case.inflow is None else dev(torch, case.inflow))
    floor = 1e-12 if case.bulk == np.float64 else 1e-6
    sens_u = relerr(Up, Uo)
    sens_q = relerr(Qp, Qo)
    e = relerr(host(Ug), Uo)  # whole array: the ghost cells of U are refilled by the stage
    eq = relerr(host(Qg), Qo)
    print(f"{case.name} fvm={p.fvm} {case.bulk} {nsteps} steps: U err {e:.2e} (sensitivity {sens_u:.2e}), "
          f"Q err {eq:.2e} (sensitivity {sens_q:.2e})")
    assert e <= 10 * sens_u + floor, f"U after {nsteps} steps: {e}"
    assert eq <= 10 * sens_q + floor, f"Q after {nsteps} steps: {eq}"


def test_stage_entry_matches_time_step(torch_mod):
    """Three f3d_rk3_stage_fvm calls equal one f3d_time_step_fvm step (same kernels, in-place Q)."""
    torch = torch_mod
    case = ramp_case(24, 16, 8, jump=1.0, fvm=f3d.FVM_EIGEN)
    p, o, ctx = case.params, Oracle(case.params), f3d.Flame3D(case.params)
    U = case.U()
    dt = 0.2 * o.cfl_dt(case.Q, case.metrics, case.J, 1.0)
    mg = [dev(torch, a) for a in case.metrics]
    Jg, Xg, ig = dev(torch, case.J), geo_dev(torch, ctx, case), dev(torch, case.inflow)
    Ua, Una, Qa = dev(torch, U), dev(torch, np.zeros_like(U)), dev(torch, case.Q)
    for KRK in (1, 2, 3):
        ctx.rk3_stage_fvm(KRK, Ua, Una, Qa, mg, Jg, Xg, dt, ig)
    Ub, Unb, Qb = dev(torch, U), dev(torch, np.zeros_like(U)), dev(torch, case.Q)
    ctx.time_step_fvm(1, Ub, Unb, Qb, mg, Jg, Xg, dt, ig)
    # heavy-face fluxes are summed with atomics, so the two runs differ in the last bits of the first
    # stage, which the later stages' sensor and WENO weights amplify (this case's response to a 1e-16
    # relative perturbation over one step is ~6e-10, test_time_step_fvm)
    e, eq = relerr(host(Ua), host(Ub)), relerr(host(Qa), host(Qb))
    print(f"stage entry vs time step: U {e:.2e}, Q {eq:.2e}")
    assert e < 1e-9 and eq < 1e-9


def test_branch_selection_is_enforced(torch_mod):
    torch = torch_mod
    fd = ramp_case(24, 16, 8)
    fv = ramp_case(24, 16, 8, fvm=f3d.FVM_CONSER)
    R = torch.zeros(fd.params.cell_shape(5), dtype=torch.float64, device="cuda:0")
    mg = [dev(torch, a) for a in fd.metrics]
    with pytest.raises(f3d.F3DError, match="finite-volume"):
        f3d.Flame3D(fv.params).rhs(R, dev(torch, fv.Q), mg, dev(torch, fv.J))
    with pytest.raises(f3d.F3DError, match="finite-difference"):
        f3d.Flame3D(fd.params).rhs_fvm(R, dev(torch, fd.Q), dev(torch, fd.U()), mg, dev(torch, fd.J),
                                       geo_dev(torch, f3d.Flame3D(fv.params), fv))
