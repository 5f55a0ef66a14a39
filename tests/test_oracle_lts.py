"""Pins of the oracle's local time stepping (params lts): compute_dt (utils.jl:73-100), div_LTS
(div.jl:52-101) and the RK3 loop with them (solver.jl:419-428).

  * compute_dt on affine meshes x = A (i, j, k): (1/J)/S_d is the spacing of the xi_d planes,
    1/|row d of A^-1|, so dt = LTS_CFL min_d(h_d/(u_d + c)) in closed form, with u_d the Cartesian
    component u, v, w as the reference writes it (not the velocity along xi_d) and u + c, not |u| + c
    (a supersonic cell moving towards -x gets a negative dt: the quirk is kept and pinned).
  * div_LTS with a uniform dt field is div with that dt, bit for bit; with any field its increment is
    the unit-dt increment scaled cell by cell.
  * A step with lts from a state of uniform velocity and temperature (so a uniform dt field) but
    varying density equals the global-dt step with that dt, for both branches; on a general state a
    step equals a loop built from compute_dt of the first stage's state and the RHS of each stage.
"""
import importlib

import numpy as np
import pytest

from casegen import ramp_case, relerr, tgv_case
from oracle_lib import Oracle

f3d = importlib.import_module("flame3d-mhd_amd")
NG = 4


def _affine(p, A, rng, dtype=np.float64):
    nx, ny, nz = p.tot
    detA = np.linalg.det(A)
    Minv = np.linalg.inv(A) * detA
    m = [np.ascontiguousarray(np.full(p.cell_shape(), Minv[a, b], dtype=dtype)) for a in range(3) for b in range(3)]
    S = [np.ascontiguousarray(np.full(p.cell_shape(), np.sqrt(np.sum(Minv[d] ** 2)), dtype=dtype)) for d in range(3)]
    J = np.full(p.cell_shape(), 1.0 / detA)
    Q = np.zeros(p.cell_shape(6))
    Q[0] = 1.0 + rng.random(Q[0].shape)
    Q[5] = 200.0 + 200.0 * rng.random(Q[0].shape)
    c = np.sqrt(p.gamma * p.Rg * Q[5])
    Q[1:4] = (2 * rng.random((3,) + Q[0].shape) - 1) * 0.9 * c
    Q[4] = Q[0] * p.Rg * Q[5]
    return np.ascontiguousarray(Q.astype(dtype)), m, J, S, np.linalg.inv(A)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sheared", [False, True])
def test_compute_dt_closed_form(dtype, sheared):
    p = f3d.make_params(7, 6, 5, lts=1, lts_cfl=0.7)
    rng = np.random.default_rng(4)
    A = np.diag([2e-3, 1e-3, 3e-3])
    if sheared:
        A = A + np.array([[0, 0.4e-3, 0.1e-3], [0.2e-3, 0, 0.3e-3], [0, 0.5e-3, 0]])
    Q, m, J, S, Ainv = _affine(p, A, rng, dtype)
    dt = np.full(p.cell_shape(), -7.0)
    Oracle(p).compute_dt(dt, Q, J, *S)
    h = 1.0 / np.sqrt(np.sum(Ainv ** 2, axis=1))  # spacing of the xi_d planes
    q = Q.astype(np.float64)
    c = np.sqrt(p.gamma * p.Rg * q[5])
    ref = 0.7 * np.minimum(np.minimum(h[0] / (q[1] + c), h[1] / (q[2] + c)), h[2] / (q[3] + c))
    inner = (slice(NG, -NG),) * 3
    tol = 1e-14 if dtype == np.float64 else 1e-6  # float: S is the float-rounded norm
    assert np.max(np.abs(dt[inner] - ref[inner]) / np.abs(ref[inner])) < tol
    ghost = np.ones(p.cell_shape(), bool)
    ghost[inner] = False
    assert np.all(dt[ghost] == -7.0), "ghost cells of dt written"


def test_compute_dt_keeps_u_plus_c():
    """u + c as written: a cell with u = -2c along x gets dt = cfl * h/(u + c) < 0 (|u| + c would give
    cfl * min_d h_d/(|u_d| + c) > 0)."""
    p = f3d.make_params(6, 5, 4, lts=1, lts_cfl=0.5)
    rng = np.random.default_rng(2)
    h = 1e-3
    Q, m, J, S, _ = _affine(p, np.diag([h, h, h]), rng)
    c = np.sqrt(p.gamma * p.Rg * Q[5])
    Q[1] = -2.0 * c
    Q[2] = 0.1 * c
    Q[3] = 0.0
    dt = np.zeros(p.cell_shape())
    Oracle(p).compute_dt(dt, np.ascontiguousarray(Q), J, *S)
    inner = (slice(NG, -NG),) * 3
    assert np.all(dt[inner] < 0)
    assert np.allclose(dt[inner], 0.5 * h / (-c[inner]), rtol=1e-14, atol=0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_div_lts_against_div(dtype):
    p = f3d.make_params(8, 6, 5)
    o = Oracle(p)
    rng = np.random.default_rng(9)
    F = [rng.standard_normal(p.face_shape(d)).astype(dtype) for d in range(3)]
    Fv = [rng.standard_normal(p.fv_shape()).astype(dtype) for _ in range(3)]
    J = 1.0 + rng.random(p.cell_shape())
    U = rng.standard_normal(p.cell_shape(5))
    # uniform field: bit for bit the scalar-dt div
    dt0 = 3.7e-4
    Ua, Ub = U.copy(), U.copy()
    o.div(Ua, *F, *Fv, dt0, J)
    o.div_lts(Ub, *F, *Fv, np.full(p.cell_shape(), dt0), J)
    assert np.array_equal(Ua, Ub)
    # any field: the unit-dt increment scaled cell by cell
    # (from U = 0, so the increments carry no cancellation against U)
    dtf = 1e-4 * (1 + rng.random(p.cell_shape()))
    inc1, incl = np.zeros_like(U), np.zeros_like(U)
    o.div(inc1, *F, *Fv, 1.0, J)
    o.div_lts(incl, *F, *Fv, dtf, J)
    inner = (slice(None),) + (slice(NG, -NG),) * 3
    assert relerr(incl[inner], (inc1 * dtf[None])[inner]) < 1e-15
    Ul = U.copy()
    o.div_lts(Ul, *F, *Fv, dtf, J)
    ghost = np.ones(U.shape, bool)
    ghost[inner] = False
    assert np.array_equal(Ul[ghost], U[ghost])


def _uniform_flow(case):
    """The case's density kept, velocity and temperature made uniform (ghosts included): a uniform
    local dt field but a non-trivial RHS (pressure = rho Rg T varies)."""
    Q = case.Q.astype(np.float64).copy()
    Q[1], Q[2], Q[3], Q[5] = 60.0, -15.0, 8.0, 290.0
    Q[4] = Q[0] * case.params.Rg * Q[5]
    case.Q = np.ascontiguousarray(Q.astype(case.Q.dtype))
    return case


@pytest.mark.parametrize("branch", ["fdm", "fvm"])
def test_lts_step_with_uniform_dt_is_the_global_step(branch):
    kw = dict(fvm=f3d.FVM_EIGEN) if branch == "fvm" else {}
    case = _uniform_flow(tgv_case(12, **kw, lts=1, lts_cfl=0.3))
    p = case.params
    o = Oracle(p)
    S = [np.ascontiguousarray(s) for s in case.s]
    dtf = np.zeros(p.cell_shape())
    o.compute_dt(dtf, case.Q, case.J, *S)
    inner = (slice(NG, -NG),) * 3
    d0 = float(dtf[inner].flat[0])
    assert d0 > 0 and np.max(np.abs(dtf[inner] - d0)) <= 4e-16 * d0
    glob = type(p).from_buffer_copy(p)
    glob.lts = 0
    og = Oracle(glob)
    U = case.U()
    Ul, Ug = U.copy(), U.copy()
    Ql, Qg = case.Q.copy(), case.Q.copy()
    # one step: its three stages use the dt field of its first stage's (uniform) state
    if branch == "fvm":
        o.time_step_fvm(1, Ul, np.zeros_like(U), Ql, case.metrics, case.J, case.coords, 123.0, case.inflow)
        og.time_step_fvm(1, Ug, np.zeros_like(U), Qg, case.metrics, case.J, case.coords, d0, case.inflow)
    else:
        o.time_step(1, Ul, np.zeros_like(U), Ql, case.metrics, case.J, 123.0, case.inflow)
        og.time_step(1, Ug, np.zeros_like(U), Qg, case.metrics, case.J, d0, case.inflow)
    # the scalar dt (123) is ignored under lts; the step moved the state
    assert relerr(Ug, U) > 1e-6
    assert relerr(Ul, Ug) < 1e-13, relerr(Ul, Ug)


def test_lts_step_uses_the_stage1_dt_of_each_step():
    """On a non-uniform state the lts step differs from the global step of any single dt, and equals
    a hand-made loop: stage 1's state sets the per-cell dt, div_LTS in all three stages."""
    case = ramp_case(16, 12, 8, lts=1, lts_cfl=0.4)
    p = case.params
    o = Oracle(p)
    U = case.U()
    Ul, Ql = U.copy(), case.Q.copy()
    o.time_step(1, Ul, np.zeros_like(U), Ql, case.metrics, case.J, 0.0, case.inflow)
    # hand loop from the kernel-level pieces
    S = [np.ascontiguousarray(s) for s in case.s]
    dtf = np.zeros(p.cell_shape())
    o.compute_dt(dtf, case.Q, case.J, *S)
    glob = type(p).from_buffer_copy(p)
    glob.lts = 0
    og = Oracle(glob)
    Uh, Un, Qh = U.copy(), U.copy(), case.Q.copy()
    for KRK in (1, 2, 3):
        R = np.zeros(p.cell_shape(5))
        og.rhs(R, Qh, case.metrics, case.J)
        Uh[:, NG:-NG, NG:-NG, NG:-NG] += (R * dtf[None])[:, NG:-NG, NG:-NG, NG:-NG]
        if KRK == 2:
            Uh = Uh * 0.25 + Un * 0.75
        elif KRK == 3:
            Uh = Uh * (2.0 / 3.0) + Un * (1.0 / 3.0)
        Uh = np.where(np.broadcast_to(_interior_mask(p), Uh.shape), Uh, U)
        og.c2prim(Uh, Qh)
        og.exchange(Qh, 6)
        og.fill(Qh, Uh, case.inflow)
    inner = (slice(None),) + (slice(NG, -NG),) * 3
    # the hand loop rounds R*dt where the stage rounds s*(J*dt); WENO weights amplify that to ~1e-11,
    # while recomputing dt in later stages or using the scalar dt moves the state by > 1e-4
    assert relerr(Ul[inner], Uh[inner]) < 1e-9, relerr(Ul[inner], Uh[inner])


def _interior_mask(p):
    m = np.zeros(p.cell_shape(), bool)
    m[NG:-NG, NG:-NG, NG:-NG] = True
    return m
