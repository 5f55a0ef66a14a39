"""f3d_rhs_noalloc_f64 (the FP64 RHS with the library's static per-GPU scratch instead of a workspace)
against f3d_rhs_f64 through the C-ABI: two calls in a row, the second on a smaller grid (the scratch
keeps the first call's contents), each against the workspace entry on the same input and against the
oracle's RHS (the FP64 parity bar)."""
import numpy as np
import pytest

from casegen import f3d, fp64_rhs_check, ramp_case, rhs_scales, scaled_err, tgv_case

pytestmark = pytest.mark.gpu


def _run(torch, case, with_phi):
    p = case.params
    ctx = f3d.Flame3D(p)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
    Q, m, J = d(case.Q), [d(a) for a in case.metrics], d(case.J)
    R_na = torch.zeros(p.cell_shape(5), dtype=torch.float64, device="cuda:0")
    R_ws = torch.zeros_like(R_na)
    phi_na = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0") if with_phi else None
    phi_ws = torch.zeros(p.cell_shape(), dtype=torch.float64, device="cuda:0")
    ctx.rhs_noalloc(R_na, Q, m, J, phi_na)
    ctx.rhs(R_ws, Q, m, J, phi_ws)
    torch.cuda.synchronize()
    return R_na.cpu().numpy(), R_ws.cpu().numpy(), (None if phi_na is None else phi_na.cpu().numpy()), \
        phi_ws.cpu().numpy()


def test_rhs_noalloc_matches_workspace_entry():
    import torch
    assert torch.cuda.is_available(), "GPU test"
    # first call: ramp with a jump and spikes (faces on WENO7, WENO5 and NND), phi from the scratch
    big = ramp_case(40, 24, 16, jump=1.0, spikes=[(5, 3, 2, 20.0), (25, 10, 7, 3.0)], hybrid_phi3=2.5)
    # second call: a smaller grid with the caller's phi
    small = tgv_case(12)
    for case, with_phi in ((big, False), (small, True)):
        print(f"no-alloc RHS: {case.name} phi {'caller' if with_phi else 'scratch'}", flush=True)
        R_na, R_ws, phi_na, phi_ws = _run(torch, case, with_phi)
        sc = rhs_scales(case)
        e = scaled_err(R_na, R_ws, sc)
        print(f"{case.name}: no-alloc vs workspace entry {e:.2e} (term-scaled)")
        # the same kernels on the same input; the heavy-face corrections may add in another order
        assert e <= 1e-14, e
        if with_phi:
            assert np.array_equal(phi_na, phi_ws)
        fp64_rhs_check(f"no-alloc {case.name}", R_na, case)


def test_rhs_noalloc_rejects_what_it_cannot_hold():
    import torch
    assert torch.cuda.is_available(), "GPU test"
    p = f3d.make_params(128, 128, 128)
    ctx = f3d.Flame3D(p)
    z = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    with pytest.raises(f3d.F3DError, match="scratch"):
        ctx.rhs_noalloc(z, z, [z] * 9, z)
    p2 = f3d.make_params(12, 12, 12, les=1)
    with pytest.raises(f3d.F3DError, match="fused engine"):
        f3d.Flame3D(p2).rhs_noalloc(z, z, [z] * 9, z)
