"""Run-to-run determinism of the fused stage: the heavy-face corrections are added one cell at a time in a
fixed order (no floating-point atomics), so repeated time steps from the same input give the same bits,
in both numeric forms and both branches, on inputs with faces on WENO7, WENO5 and NND, whatever the
context ran before."""
import numpy as np
import pytest

from casegen import f3d, ramp_case

pytestmark = pytest.mark.gpu


def _steps(torch, case, nsteps, geo=None):
    p = case.params
    ctx = f3d.Flame3D(p)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    U, Q = d(case.U()), d(case.Q)
    Un = torch.zeros_like(U)
    m = [d(a) for a in case.metrics]
    dt = 1e-9
    if geo is None:
        ctx.time_step(nsteps, U, Un, Q, m, d(case.J), dt, d(case.inflow))
    else:
        ctx.time_step_fvm(nsteps, U, Un, Q, m, d(case.J), geo, dt, d(case.inflow))
    torch.cuda.synchronize()
    return U.cpu().numpy(), Q.cpu().numpy()


@pytest.mark.parametrize("fvm", [0, 1], ids=["fdm", "fvm"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32b"])
def test_time_steps_bit_identical(fvm, dtype):
    import torch
    assert torch.cuda.is_available(), "GPU test"
    case = ramp_case(40, 24, 16, dtype=dtype, jump=1.0, spikes=[(5, 3, 2, 20.0), (25, 10, 7, 3.0)], hybrid_phi3=2.5,
                     interior_noise=0.02, fvm=f3d.FVM_EIGEN if fvm else 0)
    geo = None
    if fvm:
        p = case.params
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        geo = [torch.zeros((4,) + p.face_shape(k)[1:], dtype=tdt, device="cuda:0") for k in range(3)]
        X = [torch.from_numpy(a).to("cuda:0") for a in case.coords]
        f3d.Flame3D(p).fvm_geometry(*X, geo)
    U1, Q1 = _steps(torch, case, 3, geo)
    U2, Q2 = _steps(torch, case, 3, geo)
    print(f"fvm={fvm} {np.dtype(dtype).name}: max |U1 - U2| = {np.max(np.abs(U1 - U2)):.1e}")
    assert np.array_equal(U1, U2) and np.array_equal(Q1, Q2)


@pytest.mark.parametrize("fvm", [0, 1], ids=["fdm", "fvm"])
def test_heavy_marks_history_independent(fvm):
    """The heavy faces' cell marks are set by the forming pass and cleared by the apply pass of the same
    evaluation, so no evaluation sees another's marks: 90 steps (270 evaluations) from one state in a fresh
    context, against the same 90 steps in a context that first ran 40 steps of a state with its non-UP7 faces
    elsewhere (where stale marks would sit on other cells): the same bits.  The
    finite-difference run is also held to the oracle's 90 steps at the time-step parity bar; the finite-volume
    one is not, since on this input (a 20x pressure spike) the FP64 oracle's own finite-volume RHS is off
    its long-double evaluation by up to 1.5e-8 per variable, where the GPU's is within 1e-13."""
    import torch
    from casegen import relerr
    from oracle_lib import Oracle
    assert torch.cuda.is_available(), "GPU test"
    kw = dict(jump=1.0, hybrid_phi3=2.5, interior_noise=0.02, fvm=f3d.FVM_EIGEN if fvm else 0)
    case = ramp_case(40, 24, 16, spikes=[(5, 3, 2, 20.0), (25, 10, 7, 3.0)], **kw)
    other = ramp_case(40, 24, 16, spikes=[(30, 18, 12, 20.0), (12, 20, 3, 3.0)], **kw)
    p = case.params
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    geo = None
    if fvm:
        geo = [torch.zeros((4,) + p.face_shape(k)[1:], dtype=torch.float64, device="cuda:0") for k in range(3)]
        f3d.Flame3D(p).fvm_geometry(*[d(a) for a in case.coords], geo)
    m, J, inflow, dt = [d(a) for a in case.metrics], d(case.J), d(case.inflow), 1e-9

    def steps(ctx, c, n):
        U, Q = d(c.U()), d(c.Q)
        Un = torch.zeros_like(U)
        if geo is None:
            ctx.time_step(n, U, Un, Q, m, J, dt, inflow)
        else:
            ctx.time_step_fvm(n, U, Un, Q, m, J, geo, dt, inflow)
        torch.cuda.synchronize()
        return U.cpu().numpy(), Q.cpu().numpy()

    print(f"{'fvm' if fvm else 'fdm'}: 90 steps in a fresh context", flush=True)
    Ua, Qa = steps(f3d.Flame3D(p), case, 90)
    ctx_b = f3d.Flame3D(p)
    print("  40 steps of another state, then the same 90 steps", flush=True)
    steps(ctx_b, other, 40)
    Ub, Qb = steps(ctx_b, case, 90)
    print(f"  max |Ua - Ub| = {np.max(np.abs(Ua - Ub)):.1e}")
    assert np.array_equal(Ua, Ub) and np.array_equal(Qa, Qb)
    if fvm:
        return
    o = Oracle(p)
    Uo, Qo = case.U(), case.Q.copy()
    Up = Uo * (1 + 1e-16 * (2 * np.random.default_rng(5).random(Uo.shape) - 1))
    Qp = Qo.copy()
    o.time_step(90, Uo, np.zeros_like(Uo), Qo, case.metrics, case.J, dt, case.inflow)
    o.time_step(90, Up, np.zeros_like(Up), Qp, case.metrics, case.J, dt, case.inflow)
    inner = (slice(None),) + (slice(4, -4),) * 3
    e, sens = relerr(Ua[inner], Uo[inner]), relerr(Up[inner], Uo[inner])
    print(f"  U against the oracle: {e:.2e} (oracle sensitivity {sens:.2e})")
    assert e <= 10 * sens + 1e-12, e
