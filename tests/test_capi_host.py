"""C-ABI library loads, exports every symbol include/f3d.h declares, and validates parameters.

Only host-side functions are called (no GPU needed): version, params init/check, cell count,
last error, the ctypes struct layout, and the host mirror's shape helpers.
"""
import ctypes
import os
import re
import subprocess

import pytest

from casegen import f3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "f3d.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(f3d_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 45
    out = subprocess.run(["nm", "-D", "--defined-only", f3d.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s(f3d_\w+)", out.stdout))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    L = f3d.lib()
    for n in names:
        assert getattr(L, n) is not None


def test_struct_layout_matches_header():
    # 12 int32, 19 doubles, 8 int32, 3 doubles: the ctypes mirror must have the C size
    assert ctypes.sizeof(f3d.F3DParams) == 12 * 4 + 19 * 8 + 8 * 4 + 3 * 8
    assert f3d.F3DParams.gamma.offset == 48
    assert f3d.F3DParams.split.offset == 200 and f3d.F3DParams.lts_cfl.offset == 232


def test_params_defaults_are_the_reference_constants():
    p = f3d.make_params(64, 32, 16)
    assert (p.Nxp, p.Nyp, p.Nzp, p.NG) == (64, 32, 16, 4)
    assert (p.gamma, p.Rg, p.C_s, p.T_s, p.Pr) == (1.4, 287.0, 1.458e-6, 110.4, 0.72)   # run.jl:8-13
    assert p.Cp == pytest.approx(287.0 * 1.4 / 0.4)
    assert (p.hybrid_phi1, p.hybrid_phi2, p.hybrid_phi3, p.Linear_phi) == (0.05, 1.0, 10.0, 1.0)  # run.jl:70-73
    assert (p.weno7_eps, p.weno5_eps) == (1e-20, 1e-16)  # schemes.jl:712-713
    assert list(p.bc) == [f3d.BC_DIRICHLET, f3d.BC_ZEROGRAD, f3d.BC_WALL_ADIABATIC, f3d.BC_DIRICHLET,
                          f3d.BC_PERIODIC, f3d.BC_PERIODIC]  # boundary.jl + run.jl:19
    assert list(p.q_inf)[:5] == [1.0, 0.5, 0.0, 0.0, 1.0 / 1.4]  # boundary.jl:14-18
    assert p.q_inf[5] == pytest.approx((1.0 / 1.4) / 287.0)
    assert f3d.lib().f3d_cell_count(ctypes.byref(p)) == 72 * 40 * 24
    assert f3d.lib().f3d_version() == 10000


@pytest.mark.parametrize("field,value,code", [
    ("NG", 3, -3), ("Nxp", 2, -2), ("gamma", 1.0, -4),
])
def test_params_check_rejects(field, value, code):
    p = f3d.F3DParams()
    f3d.lib().f3d_params_init(ctypes.byref(p), 16, 16, 16)
    setattr(p, field, value)
    assert f3d.lib().f3d_params_check(ctypes.byref(p)) == code
    assert f3d.lib().f3d_last_error()


def test_params_check_bc_rules():
    L = f3d.lib()
    p = f3d.F3DParams()
    L.f3d_params_init(ctypes.byref(p), 16, 16, 16)
    p.bc[2] = 9
    assert L.f3d_params_check(ctypes.byref(p)) == -4
    L.f3d_params_init(ctypes.byref(p), 16, 16, 16)
    p.bc[3] = f3d.BC_INFLOW  # inflow only on x-low
    assert L.f3d_params_check(ctypes.byref(p)) == -4
    L.f3d_params_init(ctypes.byref(p), 16, 16, 16)
    p.bc[4] = f3d.BC_DIRICHLET  # periodic must come in pairs
    assert L.f3d_params_check(ctypes.byref(p)) == -4
    L.f3d_params_init(ctypes.byref(p), 16, 16, 16)
    p.bc[2] = f3d.BC_WALL_ISOTHERMAL  # needs T_wall
    assert L.f3d_params_check(ctypes.byref(p)) == -4
    p.T_wall = 300.0
    assert L.f3d_params_check(ctypes.byref(p)) == 0
    with pytest.raises(f3d.F3DError):
        f3d.make_params(16, 16, 16, NG=5)


def test_null_pointer_is_an_error_not_a_crash():
    p = f3d.make_params(16, 16, 16)
    L = f3d.lib()
    rc = L.f3d_shock_sensor_f64(ctypes.byref(p), None, None, None)
    assert rc == -1 and b"NULL" in L.f3d_last_error()


def test_shapes_follow_julia_column_major():
    p = f3d.make_params(10, 7, 5)
    assert p.cell_shape() == (13, 15, 18)
    assert p.cell_shape(6) == (6, 13, 15, 18)
    assert p.face_shape(0) == (5, 5, 7, 11) and p.face_shape(2) == (5, 6, 7, 10)
    assert p.fv_shape() == (4, 9, 11, 14)
    assert p.mask_shape(1) == (5, 8, 10)


def test_adjoint_entry_points_validate_before_any_launch():
    p = f3d.make_params(16, 12, 8)
    L = f3d.lib()
    # NULL adjoint workspace and NULL arrays are errors, not crashes (no GPU call is made)
    assert L.f3d_rhs_vjp_f64(ctypes.byref(p), None, None, None, None, None, None, None) == -1
    assert L.f3d_rk3_stage_vjp_f64(ctypes.byref(p), None, 1, None, None, None, None, None, None, None,
                                   ctypes.c_double(1.0), None, None, None) == -1
    assert L.f3d_ghost_vjp_f64(ctypes.byref(p), None, None, None, None, None) == -1
    assert L.f3d_c2prim_vjp_f64(ctypes.byref(p), None, None, None, None) == -1
    assert b"NULL" in L.f3d_last_error()
    # checkpoint buffer size: ceil(n/every) + every states of 5 variables, plus two stage states per
    # step of a segment (of every step when every == 1, which also keeps three primitive states of 6
    # variables per step); an interval above nsteps counts as nsteps
    nc = L.f3d_cell_count(ctypes.byref(p))
    assert L.f3d_adjoint_ckpt_bytes(ctypes.byref(p), 50, 1) == ((50 + 1 + 2 * 50) * 5 + 50 * 18) * nc * 8
    assert L.f3d_adjoint_ckpt_bytes(ctypes.byref(p), 50, 10) == (5 + 10 + 2 * 10) * 5 * nc * 8
    assert L.f3d_adjoint_ckpt_bytes(ctypes.byref(p), 7, 10) == (1 + 7 + 2 * 7) * 5 * nc * 8
    assert L.f3d_adjoint_ckpt_bytes(ctypes.byref(p), 0, 1) == 0


def test_fvm_selector_and_entry_validation():
    L = f3d.lib()
    assert f3d.make_params(16, 12, 8).fvm == f3d.FDM  # finite-difference branch unless asked
    p = f3d.make_params(16, 12, 8, fvm=f3d.FVM_EIGEN)
    q = f3d.F3DParams()
    L.f3d_params_init(ctypes.byref(q), 16, 16, 16)
    q.fvm = 3
    assert L.f3d_params_check(ctypes.byref(q)) == -4 and b"fvm" in L.f3d_last_error()
    for name in ("eigen_reconstruct_i", "eigen_reconstruct_k", "conser_reconstruct_j"):
        for suf in ("f64", "f32b"):
            fn = getattr(L, f"f3d_{name}_{suf}")
            assert fn(ctypes.byref(p), None, None, None, None, None, None, None, None, None) == -1
    assert L.f3d_rhs_fvm_f64(ctypes.byref(p), None, None, None, None, None, None, None, None, None) == -1
    assert L.f3d_fvm_geometry_f32b(ctypes.byref(p), None, None, None, None, None) == -1
    assert L.f3d_time_step_fvm_f32b(ctypes.byref(p), None, 1, None, None, None, None, None, None,
                                    ctypes.c_double(1.0), None, None) == -1
    # the adjoint differentiates the finite-difference branch only: rejected before any allocation
    h = ctypes.c_void_p()
    assert L.f3d_adjoint_workspace_create(ctypes.byref(p), ctypes.byref(h)) == -4
    assert b"finite-difference" in L.f3d_last_error()


def test_alternative_flags_validation():
    """split / character / les / lts are checked by f3d_params_check; the adjoint covers the default
    Steger-Warming + characteristic + laminar + global-dt + unfiltered path only (rejected before any
    allocation)."""
    L = f3d.lib()
    q = f3d.F3DParams()
    for field, bad in (("split", 4), ("split", -1), ("character", 2), ("les", 3), ("lts", 2)):
        L.f3d_params_init(ctypes.byref(q), 16, 16, 16)
        setattr(q, field, bad)
        assert L.f3d_params_check(ctypes.byref(q)) == -4, field
    for kw in ({"split": f3d.SPLIT_LF}, {"split": f3d.SPLIT_AUSM}, {"character": 0}, {"les": f3d.LES_WALE},
               {"lts": 1}, {"filtering": 1}):
        p = f3d.make_params(16, 12, 8, **kw)
        h = ctypes.c_void_p()
        assert L.f3d_adjoint_workspace_create(ctypes.byref(p), ctypes.byref(h)) == -4, kw
        assert b"adjoint covers" in L.f3d_last_error()
    # the other splittings and the component-wise advect validate their pointers before any launch
    p = f3d.make_params(16, 12, 8)
    for name in ("flux_split_lf", "flux_split_vl", "flux_split_ausm"):
        for suf in ("f64", "f32b"):
            assert getattr(L, f"f3d_{name}_{suf}")(ctypes.byref(p), None, None, None, None, None, None, None,
                                                   None) == -1
    for d in "xyz":
        for suf in ("f64", "f32b"):
            assert getattr(L, f"f3d_advect_{d}_{suf}")(ctypes.byref(p), None, None, None, None, None,
                                                       ctypes.c_int32(5), None) == -1
    for suf in ("f64", "f32b"):
        assert getattr(L, f"f3d_compute_dt_{suf}")(ctypes.byref(p), None, None, None, None, None, None, None) == -1
        assert getattr(L, f"f3d_div_lts_{suf}")(ctypes.byref(p), None, None, None, None, None, None, None, None,
                                                None, None) == -1
    q = f3d.F3DParams()
    L.f3d_params_init(ctypes.byref(q), 16, 16, 16)
    q.lts, q.lts_cfl = 1, 0.0
    assert L.f3d_params_check(ctypes.byref(q)) == -4 and b"lts_cfl" in L.f3d_last_error()
