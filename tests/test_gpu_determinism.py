"""Run-to-run determinism of the fused stage: the heavy-face corrections are added in two ordered passes
per direction (no floating-point atomics), so repeated time steps from the same input give the same bits,
in both numeric forms and both branches, on inputs with faces on WENO7, WENO5 and NND."""
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
