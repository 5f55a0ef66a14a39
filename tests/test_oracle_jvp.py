"""Pins of the oracle's forward mode (dual numbers, oracle/f3d_dual.hpp), the reference the GPU
adjoint is checked against: central finite differences of the oracle's own FP64 functions converge
to its JVP at second order on inputs whose stencils are resolved (seeded interior variation), the
ghost procedure's JVP is exact (it is affine), and c2Prim's JVP matches FD.  CPU only.

The density/velocity/temperature perturbations leave the pressure, hence the sensor and the scheme
mask, unchanged, so the finite differences see the same frozen mask as the JVP."""
import numpy as np

from casegen import ramp_case, tgv_case
from oracle_lib import Oracle

INT = (slice(None),) + (slice(4, -4),) * 3


def _rhs_fd_errors(case, dQ, eps_list):
    p, o = case.params, Oracle(case.params)
    R = np.zeros(p.cell_shape(5))
    dR = np.zeros_like(R)
    o.rhs_jvp(R, dR, case.Q, dQ, case.metrics, case.J)
    errs = []
    for eps in eps_list:
        Rp, Rm = np.zeros_like(R), np.zeros_like(R)
        o.rhs(Rp, case.Q + eps * dQ, case.metrics, case.J)
        o.rhs(Rm, case.Q - eps * dQ, case.metrics, case.J)
        fd = (Rp - Rm) / (2 * eps)
        errs.append(max(np.abs(fd[n] - dR[n]).max() / np.abs(dR[n]).max() for n in range(5) if np.abs(dR[n]).max() > 0))
    return errs


def _dq(case, rng, vars_=(0, 1, 2, 3, 5), rel=1e-3):
    dQ = np.zeros_like(case.Q)
    for n in vars_:
        dQ[n] = rng.standard_normal(case.Q[n].shape) * np.abs(case.Q[n]) * rel
    return dQ


def test_rhs_jvp_converges_to_fd_on_weno_ramp():
    case = ramp_case(16, 12, 8, jump=1.0, spikes=[(5, 3, 2, 20.0)], hybrid_phi3=2.5, interior_noise=1e-3)
    e = _rhs_fd_errors(case, _dq(case, np.random.default_rng(0), vars_=(0, 5)), [1e-3, 1e-4])
    print("ramp rho,T:", e)
    assert e[1] < 1e-5 and e[1] < e[0] / 30          # second order: 100x per decade, 30x asked


def test_rhs_jvp_converges_to_fd_on_tgv():
    case = tgv_case(12)
    e = _rhs_fd_errors(case, _dq(case, np.random.default_rng(1), vars_=(0, 1, 2, 3, 5)), [1e-3, 1e-4])
    print("tgv:", e)
    assert e[1] < 1e-4 and e[1] < e[0] / 30


def test_ghost_jvp_is_exact():
    case = ramp_case(16, 12, 8, bc=[5, 2, 4, 1, 0, 0])
    p, o = case.params, Oracle(case.params)
    rng = np.random.default_rng(2)
    Q = case.Q.copy()
    dQ = np.zeros_like(Q)
    dQ[INT] = rng.standard_normal(Q[INT].shape) * np.abs(Q[INT]) * 1e-3
    dinfl = rng.standard_normal(case.inflow.shape) * np.abs(case.inflow) * 1e-3
    dqinf = rng.standard_normal(6) * np.abs(np.array(list(p.q_inf))) * 1e-3
    Qd, dQd = Q.copy(), dQ.copy()
    o.ghosts_jvp(Qd, dQd, case.inflow, dinfl, dqinf, 0.7)

    def ghosts(Qx, infl, qinf, Tw):
        import importlib
        f3d = importlib.import_module("flame3d-mhd_amd")
        kw = dict(bc=list(p.bc), q_inf=list(qinf), T_wall=Tw, viscous=p.viscous)
        pp = f3d.make_params(p.Nxp, p.Nyp, p.Nzp, **kw)
        oo = Oracle(pp)
        Qx = Qx.copy()
        oo.exchange(Qx, 6)
        oo.fill(Qx, None, infl)
        return Qx

    h = 1e-6
    qinf = np.array(list(p.q_inf))
    Qp = ghosts(Q + h * dQ, case.inflow + h * dinfl, qinf + h * dqinf, p.T_wall + h * 0.7)
    Qm = ghosts(Q - h * dQ, case.inflow - h * dinfl, qinf - h * dqinf, p.T_wall - h * 0.7)
    fd = (Qp - Qm) / (2 * h)
    for n in range(6):
        scale = np.abs(dQd[n]).max()
        assert np.abs(fd[n] - dQd[n]).max() <= 1e-6 * scale, n   # rho ghost of the isothermal wall is p/(Rg T)


def test_c2prim_jvp_matches_fd():
    case = ramp_case(16, 12, 8)
    p, o = case.params, Oracle(case.params)
    U = case.U()
    rng = np.random.default_rng(3)
    dU = np.zeros_like(U)
    dU[INT] = rng.standard_normal(U[INT].shape) * np.abs(U[INT]) * 1e-3
    Q, dQ = np.zeros(p.cell_shape(6)), np.zeros(p.cell_shape(6))
    o.c2prim_jvp(U, dU, Q, dQ)
    h = 1e-4
    Qp, Qm = np.zeros_like(Q), np.zeros_like(Q)
    o.c2prim(U + h * dU, Qp)
    o.c2prim(U - h * dU, Qm)
    fd = (Qp - Qm) / (2 * h)
    for n in range(6):
        assert np.abs(fd[n] - dQ[n]).max() <= 1e-6 * np.abs(dQ[n]).max(), n


def test_multistep_jvp_converges_to_fd():
    case = ramp_case(16, 12, 8, jump=0.5, interior_noise=1e-3)
    p, o = case.params, Oracle(case.params)
    U0 = case.U()
    dt = o.cfl_dt(case.Q, case.metrics, case.J, 0.3)
    rng = np.random.default_rng(4)
    dU = np.zeros_like(U0)
    dU[INT] = rng.standard_normal(U0[INT].shape) * np.abs(U0[INT]) * 1e-3
    W = np.zeros(p.cell_shape(6))
    W[4, 4:-4, 4, 4:-4] = 1.0

    def J_of(U):
        Uc, d = U.copy(), np.zeros_like(U)
        return o.time_step_jvp(2, Uc, d, case.metrics, case.J, dt, inflow=case.inflow, W=W)[0]

    Uc, d = U0.copy(), dU.copy()
    _, dJ = o.time_step_jvp(2, Uc, d, case.metrics, case.J, dt, inflow=case.inflow, W=W)
    errs = []
    for eps in (1e-2, 1e-3):
        fd = (J_of(U0 + eps * dU) - J_of(U0 - eps * dU)) / (2 * eps)
        errs.append(abs(fd - dJ) / abs(dJ))
    print("multi-step:", errs)
    assert errs[1] < 1e-5 and errs[1] < errs[0] / 30
