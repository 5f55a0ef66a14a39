"""ctypes wrapper of the CPU oracle (oracle/liboracle_f3d.so).  Test infrastructure only.

Arrays are numpy, C-order with the reversed (Julia column-major) shape used by the package.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# F3D_ORACLE_SO: another build of the oracle (a copy with a planted mistake, to check that the pins fail)
ORACLE_SO = os.environ.get("F3D_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle_f3d.so")

_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return ORACLE_SO


def olib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        L.oracle_cfl_dt_f64.restype = c_double
        L.oracle_cfl_dt_f32b.restype = c_double
        _lib = L
    return _lib


def P(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return c_void_p(a.ctypes.data)


def PP(arrs):
    arr = (c_void_p * len(arrs))()
    for i, a in enumerate(arrs):
        assert a.flags["C_CONTIGUOUS"]
        arr[i] = a.ctypes.data
    return arr


def suf(a):
    return "f64" if a.dtype == np.float64 else "f32b"


class Oracle:
    """Oracle calls bound to one params struct (a flame3d F3DParams instance)."""

    def __init__(self, params):
        self.p = params
        self.L = olib()

    def _c(self, name, *args):
        getattr(self.L, name)(ctypes.byref(self.p), *args)

    def shock_sensor(self, phi, Q):
        self._c(f"oracle_shock_sensor_{suf(Q)}", P(phi), P(Q))

    def scheme_mask(self, mx, my, mz, phi):
        self._c(f"oracle_scheme_mask_{suf(phi)}", P(mx), P(my), P(mz), P(phi))

    def flux_split(self, Q, Fp, Fm, S, Ax, Ay, Az):
        self._c(f"oracle_flux_split_sw_{suf(Q)}", P(Q), P(Fp), P(Fm), P(S), P(Ax), P(Ay), P(Az))

    def flux_split_method(self, method, Q, Fp, Fm, S, Ax, Ay, Az):
        """fluxSplit_SW / _LF / _VL / _AUSM by method index (F3D_SPLIT_*)."""
        self._c(f"oracle_flux_split_{suf(Q)}", c_int(method), P(Q), P(Fp), P(Fm), P(S), P(Ax), P(Ay), P(Az))

    def advect_comp(self, d, F, phi, S, Fp, Fm, NV=5):
        """Component-wise advect_x / _y / _z (schemes.jl:86-698)."""
        self._c(f"oracle_advect_{suf(Fp)}", c_int(d), P(F), P(phi), P(S), P(Fp), P(Fm), c_int(NV))

    def advect(self, d, F, phi, S, Fp, Fm, Q, Ax, Ay, Az):
        self._c(f"oracle_advect_c_{suf(Q)}", c_int(d), P(F), P(phi), P(S), P(Fp), P(Fm), P(Q), P(Ax), P(Ay), P(Az))

    def compute_dt(self, dt, Q, J, S1, S2, S3):
        """compute_dt (utils.jl:73-100): per-cell local time step of the interior (dt double)."""
        self._c(f"oracle_compute_dt_{suf(Q)}", P(dt), P(Q), P(J), P(S1), P(S2), P(S3))

    def div_lts(self, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J):
        """div_LTS (div.jl:52-101) with the per-cell dt array."""
        self._c(f"oracle_div_lts_{suf(Fx)}", P(U), P(Fx), P(Fy), P(Fz), P(Fvx), P(Fvy), P(Fvz), P(dt), P(J))

    def pre_sensor(self, d, sc, Q, rth):
        """pre_x/_y/_z (utils.jl:102-169) into the cell-shaped double sc."""
        self._c(f"oracle_pre_sensor_{suf(Q)}", c_int(d), P(sc), P(Q), c_double(rth))

    def filter_nonlinear(self, d, U, Un, sc, s0):
        """filter_x/_y/_z (utils.jl:171-241)."""
        self._c("oracle_filter_nonlinear", c_int(d), P(U), P(Un), P(sc), c_double(s0))

    def filter_linear(self, d, U, Un, s0):
        """linearFilter_x/_y/_z (utils.jl:243-316)."""
        self._c("oracle_filter_linear", c_int(d), P(U), P(Un), c_double(s0))

    def filter_block(self, U, Un, Q, sc):
        """The filtering block of the time loop (solver.jl:446-470) with the params' filter settings."""
        self._c(f"oracle_filter_block_{suf(Q)}", P(U), P(Un), P(Q), P(sc))

    def viscous(self, Fvx, Fvy, Fvz, Q, m, J):
        self._c(f"oracle_viscous_flux_{suf(Q)}", P(Fvx), P(Fvy), P(Fvz), P(Q), PP(m), P(J))

    def div(self, U, Fx, Fy, Fz, Fvx, Fvy, Fvz, dt, J):
        self._c(f"oracle_div_{suf(Fx)}", P(U), P(Fx), P(Fy), P(Fz), P(Fvx), P(Fvy), P(Fvz), c_double(dt), P(J))

    def lin_comb(self, U, Un, NV, a, b):
        self._c("oracle_lin_comb", P(U), P(Un), c_int(NV), c_double(a), c_double(b))

    def c2prim(self, U, Q):
        self._c(f"oracle_c2prim_{suf(Q)}", P(U), P(Q))

    def prim2c(self, U, Q):
        self._c(f"oracle_prim2c_{suf(Q)}", P(U), P(Q))

    def exchange(self, Q, NV):
        self._c(f"oracle_exchange_ghost_{suf(Q)}", P(Q), c_int(NV))

    def fill(self, Q, U, inflow=None):
        self._c(f"oracle_fill_ghost_{suf(Q)}", P(Q), P(U), P(inflow))

    def norms(self, m, s1, s2, s3):
        self._c(f"oracle_metric_norms_{suf(s1)}", PP(m), P(s1), P(s2), P(s3))

    def rhs(self, R, Q, m, J, phi=None, masks=(None, None, None)):
        self._c(f"oracle_rhs_{suf(Q)}", P(R), P(Q), PP(m), P(J), P(phi), *map(P, masks))

    def rhs_absterms(self, A, Q, m, J):
        """J * sum of |terms| of the RHS per cell and variable (scale of the parity bar)."""
        self._c(f"oracle_rhs_absterms_{suf(Q)}", P(A), P(Q), PP(m), P(J))

    def rhs_jvp(self, R, dR, Q, dQ, m, J):
        """Forward-mode JVP of the FP64 RHS w.r.t. Q (all cells), mask frozen at Q."""
        self._c("oracle_rhs_jvp_f64", P(R), P(dR), P(Q), P(dQ), PP(m), P(J))

    def time_step_jvp(self, nsteps, U, dU, m, J, dt, inflow=None, dinflow=None, dqinf=None, dTw=None, W=None):
        """Forward-mode derivative of nsteps FP64 RK3 steps from Q0 = ghosts(c2Prim(U0)), mask frozen;
        U/dU are advanced in place.  Returns (objective, d objective) for sum_n dt <W, Q^{n+1}>."""
        obj = ctypes.c_double(0.0)
        dobj = ctypes.c_double(0.0)
        dq = None if dqinf is None else np.ascontiguousarray(dqinf, dtype=np.float64)
        dt_w = None if dTw is None else np.array([dTw], dtype=np.float64)
        self._c("oracle_time_step_jvp_f64", int(nsteps), P(U), P(dU), P(inflow), P(dinflow), P(dq), P(dt_w),
                PP(m), P(J), ctypes.c_double(dt), P(W), ctypes.byref(obj), ctypes.byref(dobj))
        return obj.value, dobj.value

    def c2prim_jvp(self, U, dU, Q, dQ):
        self._c("oracle_c2prim_jvp_f64", P(U), P(dU), P(Q), P(dQ))

    def ghosts_jvp(self, Q, dQ, inflow=None, dinflow=None, dqinf=None, dTw=None):
        dq = None if dqinf is None else np.ascontiguousarray(dqinf, dtype=np.float64)
        dt_w = None if dTw is None else np.array([dTw], dtype=np.float64)
        self._c("oracle_ghosts_jvp_f64", P(Q), P(dQ), P(inflow), P(dinflow), P(dq), P(dt_w))

    def rhs_ld(self, R, Q, m, J):
        """RHS in long double from float64 inputs, mask frozen from the float64 sensor."""
        self._c("oracle_rhs_ld", P(R), P(Q), PP(m), P(J))

    def advect_ld(self, d, F, phi, S, Fp, Fm, Q, Ax, Ay, Az):
        """advect_c in long double from float64 inputs (rounding-floor reference)."""
        self._c("oracle_advect_c_ld", c_int(d), P(F), P(phi), P(S), P(Fp), P(Fm), P(Q), P(Ax), P(Ay), P(Az))

    def time_step(self, nsteps, U, Un, Q, m, J, dt, inflow=None):
        self._c(f"oracle_time_step_{suf(Q)}", c_int(nsteps), P(U), P(Un), P(Q), PP(m), P(J), c_double(dt), P(inflow))

    def cfl_dt(self, Q, m, J, cfl):
        return getattr(self.L, f"oracle_cfl_dt_{suf(Q)}")(ctypes.byref(self.p), P(Q), PP(m), P(J), c_double(cfl))

    def metrics(self, x, y, z, m, J):
        self._c("oracle_metrics_f64", P(x), P(y), P(z), PP(m), P(J))

    # -- finite-volume branch (FVM.jl) --------------------------------------------------------
    def fvm_reconstruct(self, d, eigen, Q, U, phi, S, F, x, y, z, UL=None, UR=None, geo=None):
        """Eigen_/Conser_reconstruct of direction d: F = HLLC * Area; optional UL/UR states and
        geo (nx, ny, nz, Area) per face."""
        self._c(f"oracle_fvm_reconstruct_{suf(Q)}", c_int(d), c_int(1 if eigen else 0), P(Q), P(U), P(phi), P(S),
                P(F), P(x), P(y), P(z), P(UL), P(UR), P(geo))

    def fvm_reconstruct_ld(self, d, eigen, Q, U, phi, S, F, x, y, z):
        """Face fluxes in long double from float64 inputs (rounding-floor reference)."""
        self._c("oracle_fvm_reconstruct_ld", c_int(d), c_int(1 if eigen else 0), P(Q), P(U), P(phi), P(S), P(F),
                P(x), P(y), P(z))

    def fvm_geometry(self, d, x, y, z, geo):
        self._c("oracle_fvm_geometry_f64", c_int(d), P(x), P(y), P(z), P(geo))

    def hllc(self, UL, UR, n, F):
        """HLLC of independent faces: UL, UR, F (5, nf); n (3, nf); float64."""
        nf = UL.shape[1]
        self._c("oracle_hllc_f64", ctypes.c_long(nf), P(UL), P(UR), P(n), P(F))

    def rhs_fvm(self, R, Q, U, m, J, coords, phi=None):
        self._c(f"oracle_rhs_fvm_{suf(Q)}", P(R), P(Q), P(U), PP(m), P(J), PP(coords), P(phi))

    def rhs_fvm_ld(self, R, Q, U, m, J, coords):
        self._c("oracle_rhs_fvm_ld", P(R), P(Q), P(U), PP(m), P(J), PP(coords))

    def rhs_absterms_fvm(self, A, Q, U, m, J, coords):
        self._c(f"oracle_rhs_absterms_fvm_{suf(Q)}", P(A), P(Q), P(U), PP(m), P(J), PP(coords))

    def time_step_fvm(self, nsteps, U, Un, Q, m, J, coords, dt, inflow=None):
        self._c(f"oracle_time_step_fvm_{suf(Q)}", c_int(nsteps), P(U), P(Un), P(Q), PP(m), P(J), PP(coords),
                c_double(dt), P(inflow))

def max_threads():
    return olib().oracle_max_threads()


WENO_FORMS = {"f64": 0, "ld": 1, "dual": 2, "f32": 3}


def weno_scalar(kind, form, V, s0, eps=0.0):
    """One scalar WENO reconstruction of the oracle: kind 0 finite-difference WENO7 (returns 12x the
    face value), 1 finite-difference WENO5 on V[0..4] (6x), 2 finite-volume WENO7 (the value; its own
    eps), 3 finite-volume WENO5 on V[1..5] (the value; its own eps).  form: a WENO_FORMS key."""
    L = olib()
    f = L.oracle_weno_scalar
    f.restype = c_double
    f.argtypes = [c_int, c_int, c_void_p, c_double, c_double]
    v = np.zeros(7)
    v[:len(V)] = np.asarray(V, dtype=np.float64)
    return f(int(kind), WENO_FORMS[form], v.ctypes.data, float(s0), float(eps))
