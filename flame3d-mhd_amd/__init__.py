"""Flame3D RHS + RK3 stage on MI355X (gfx950): Python host over libf3d_hip.so.

The reference (rygan-code/Flame3D-MHD, Julia + CUDA.jl) launches per-stage kernels from
`time_step` (solver.jl:191-680).  This module exposes the same entry points with the same names
and argument order -- `shockSensor`, `fluxSplit_SW/LF/VL/AUSM`, `advect_xc/yc/zc`, the
component-wise `advect_x/y/z`, `viscousFlux`, `div`, `div_LTS`, `compute_dt`, the filters
`pre_x/y/z`, `filter_x/y/z`, `linearFilter_x/y/z`,
`linComb`, `c2Prim`, `prim2c`, `exchange_ghost`, `fillGhost`, `flowAdvance`, `time_step`, and the
finite-volume branch's `Eigen_reconstruct_{i,j,k}` / `Conser_reconstruct_{i,j,k}` (FVM.jl) -- as
methods of a `Flame3D` context that carries what the reference keeps in `const` globals
(run.jl:3-108).  Arrays are torch tensors on the GPU in the reference's column-major layout:
a Julia array (Nx_tot, Ny_tot, Nz_tot, nvar) is a C-contiguous tensor of shape
(nvar, Nz_tot, Ny_tot, Nx_tot).  Every call goes through the C-ABI of include/f3d.h; there is no
CPU fallback: if the library is missing, importing the module fails.

Errors: a non-zero status from the library raises `F3DError` with the library's message (the
reference raises Julia `error(...)` for bad configuration, solver.jl:143, run.jl:119);
`nan_check` raises `F3DNaNError` where the reference calls MPI.Abort (solver.jl:481-486).
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_size_t, c_void_p

__all__ = [
    "F3DParams", "F3DError", "F3DNaNError", "Flame3D", "Workspace", "AdjointWorkspace", "RHSFunction", "lib",
    "make_params",
    "BC_PERIODIC", "BC_DIRICHLET", "BC_ZEROGRAD", "BC_WALL_ADIABATIC", "BC_WALL_ISOTHERMAL", "BC_INFLOW",
    "PREC_F64", "PREC_F32B", "SCHEME_UP7", "SCHEME_WENO7", "SCHEME_WENO5", "SCHEME_NND2",
    "FDM", "FVM_EIGEN", "FVM_CONSER", "SPLIT_SW", "SPLIT_LF", "SPLIT_VL", "SPLIT_AUSM",
    "LES_NONE", "LES_SMAG", "LES_WALE",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libf3d_hip.so")

BC_PERIODIC, BC_DIRICHLET, BC_ZEROGRAD, BC_WALL_ADIABATIC, BC_WALL_ISOTHERMAL, BC_INFLOW = range(6)
PREC_F64, PREC_F32B = 0, 1
SCHEME_UP7, SCHEME_WENO7, SCHEME_WENO5, SCHEME_NND2 = range(4)
FDM, FVM_EIGEN, FVM_CONSER = range(3)
SPLIT_SW, SPLIT_LF, SPLIT_VL, SPLIT_AUSM = range(4)
LES_NONE, LES_SMAG, LES_WALE = range(3)
NG = 4


class F3DParams(ctypes.Structure):
    """ctypes mirror of `f3d_params` (include/f3d.h)."""
    _fields_ = [
        ("Nxp", c_int32), ("Nyp", c_int32), ("Nzp", c_int32), ("NG", c_int32), ("viscous", c_int32),
        ("bc", c_int32 * 6), ("fvm", c_int32),
        ("gamma", c_double), ("Rg", c_double), ("Cp", c_double), ("C_s", c_double), ("T_s", c_double),
        ("Pr", c_double), ("hybrid_phi1", c_double), ("hybrid_phi2", c_double), ("hybrid_phi3", c_double),
        ("Linear_phi", c_double), ("weno7_eps", c_double), ("weno5_eps", c_double),
        ("q_inf", c_double * 6), ("T_wall", c_double),
        ("split", c_int32), ("character", c_int32), ("les", c_int32), ("lts", c_int32),
        ("filtering", c_int32), ("filtering_nonlinear", c_int32), ("filtering_interval", c_int32), ("pad_", c_int32),
        ("lts_cfl", c_double), ("filtering_rth", c_double), ("filtering_s0", c_double),
    ]

    @property
    def dims(self):
        return self.Nxp, self.Nyp, self.Nzp

    @property
    def tot(self):
        return self.Nxp + 2 * self.NG, self.Nyp + 2 * self.NG, self.Nzp + 2 * self.NG

    def cell_shape(self, nvar=None):
        """torch/numpy shape of a Julia (Nx_tot, Ny_tot, Nz_tot[, nvar]) array."""
        nx, ny, nz = self.tot
        return (nz, ny, nx) if nvar is None else (nvar, nz, ny, nx)

    def face_shape(self, d):
        nx, ny, nz = self.Nxp + (d == 0), self.Nyp + (d == 1), self.Nzp + (d == 2)
        return (5, nz, ny, nx)

    def fv_shape(self):
        return (4, self.Nzp + self.NG, self.Nyp + self.NG, self.Nxp + self.NG)

    def mask_shape(self, d):
        return self.face_shape(d)[1:]


class F3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"f3d error {code}: {msg}")
        self.code = code


class F3DNaNError(F3DError):
    pass


_lib = None


def lib():
    """The loaded libf3d_hip.so.  Raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.f3d_last_error.restype = ctypes.c_char_p
        L.f3d_params_init.restype = None
        L.f3d_params_init.argtypes = [POINTER(F3DParams), c_int32, c_int32, c_int32]
        L.f3d_cell_count.restype = c_size_t
        L.f3d_workspace_bytes.restype = c_size_t
        L.f3d_workspace_create.argtypes = [POINTER(F3DParams), c_int32, POINTER(c_void_p)]
        L.f3d_workspace_destroy.argtypes = [c_void_p]
        L.f3d_workspace_bytes.argtypes = [c_void_p]
        L.f3d_adjoint_workspace_create.argtypes = [POINTER(F3DParams), POINTER(c_void_p)]
        L.f3d_adjoint_workspace_destroy.argtypes = [c_void_p]
        L.f3d_adjoint_workspace_bytes.argtypes = [c_void_p]
        L.f3d_adjoint_workspace_bytes.restype = c_size_t
        L.f3d_adjoint_ckpt_bytes.argtypes = [POINTER(F3DParams), c_int32, c_int32]
        L.f3d_adjoint_ckpt_bytes.restype = c_size_t
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().f3d_last_error().decode(errors="replace")
        if rc == -5:
            raise F3DNaNError(rc, msg)
        raise F3DError(rc, msg)


def make_params(Nxp, Nyp, Nzp, **kw):
    """Reference defaults (run.jl constants, boundary.jl boundary set), then overrides.

    `bc` may be given as a 6-sequence (xlo, xhi, ylo, yhi, zlo, zhi), `q_inf` as a 6-sequence.
    """
    p = F3DParams()
    lib().f3d_params_init(ctypes.byref(p), Nxp, Nyp, Nzp)
    for k, v in kw.items():
        if k in ("bc", "q_inf"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
    if "gamma" in kw or "Rg" in kw:
        if "Cp" not in kw:
            p.Cp = p.Rg * p.gamma / (p.gamma - 1)
    _check(lib().f3d_params_check(ctypes.byref(p)))
    return p


def _torch():
    import torch
    return torch


def _ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return c_void_p(t.data_ptr())


def _ptrs(ts):
    arr = (c_void_p * len(ts))()
    for i, t in enumerate(ts):
        arr[i] = t.data_ptr() if t is not None else None
    return arr


class Workspace:
    """Owner of an f3d_workspace (device scratch of the fused path)."""

    def __init__(self, params, precision):
        self.params = params
        self.precision = precision
        self._h = c_void_p()
        _check(lib().f3d_workspace_create(ctypes.byref(params), precision, ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h

    @property
    def nbytes(self):
        return lib().f3d_workspace_bytes(self._h)

    def close(self):
        if self._h:
            _check(lib().f3d_workspace_destroy(self._h))
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AdjointWorkspace:
    """Owner of an f3d_adj_workspace (device scratch of the FP64 adjoint, with its own forward
    workspace for stage recomputation)."""

    def __init__(self, params):
        self.params = params
        self._h = c_void_p()
        _check(lib().f3d_adjoint_workspace_create(ctypes.byref(params), ctypes.byref(self._h)))

    @property
    def handle(self):
        return self._h

    @property
    def nbytes(self):
        return lib().f3d_adjoint_workspace_bytes(self._h)

    def close(self):
        if self._h:
            _check(lib().f3d_adjoint_workspace_destroy(self._h))
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Flame3D:
    """Context holding the run.jl constants (`params`) and a HIP stream.

    Methods carry the reference's names and argument lists.  The numeric form is chosen from the
    tensors: all-float64 -> `_f64`; float32 bulk arrays with float64 U/Un/J -> `_f32b`.
    """

    def __init__(self, params, stream=None):
        self.params = params
        self._stream = stream
        self._ws = {}
        self._aws = None

    # -- plumbing ---------------------------------------------------------------------------
    @property
    def stream(self):
        import torch
        s = self._stream if self._stream is not None else torch.cuda.current_stream()
        return c_void_p(s.cuda_stream)

    def _suffix(self, t):
        import torch
        if t.dtype == torch.float64:
            return "f64"
        if t.dtype == torch.float32:
            return "f32b"
        raise TypeError(f"unsupported dtype {t.dtype}")

    def _call(self, name, *args):
        fn = getattr(lib(), name)
        _check(fn(ctypes.byref(self.params), *args, self.stream))

    def workspace(self, precision):
        ws = self._ws.get(precision)
        if ws is None:
            ws = self._ws[precision] = Workspace(self.params, precision)
        return ws

    # -- the reference's per-stage kernels ----------------------------------------------------
    def shockSensor(self, phi, Q):
        """schemes.jl:1-24"""
        self._call(f"f3d_shock_sensor_{self._suffix(Q)}", _ptr(phi), _ptr(Q))

    def schemeMask(self, mask_x, mask_y, mask_z, phi):
        """Face scheme index of advect_{x,y,z}c (schemes.jl:724-731, 813, 931)."""
        self._call(f"f3d_scheme_mask_{self._suffix(phi)}", _ptr(mask_x), _ptr(mask_y), _ptr(mask_z), _ptr(phi))

    def fluxSplit_SW(self, Q, Fp, Fm, S, Ax, Ay, Az):
        """split.jl:2-68"""
        self._call(f"f3d_flux_split_sw_{self._suffix(Q)}", *map(_ptr, (Q, Fp, Fm, S, Ax, Ay, Az)))

    def fluxSplit_LF(self, Q, Fp, Fm, S, Ax, Ay, Az):
        """split.jl:70-108 (local Lax-Friedrichs)"""
        self._call(f"f3d_flux_split_lf_{self._suffix(Q)}", *map(_ptr, (Q, Fp, Fm, S, Ax, Ay, Az)))

    def fluxSplit_VL(self, Q, Fp, Fm, S, Ax, Ay, Az):
        """split.jl:110-183 (van Leer)"""
        self._call(f"f3d_flux_split_vl_{self._suffix(Q)}", *map(_ptr, (Q, Fp, Fm, S, Ax, Ay, Az)))

    def fluxSplit_AUSM(self, Q, Fp, Fm, S, Ax, Ay, Az):
        """split.jl:185-255 (AUSM)"""
        self._call(f"f3d_flux_split_ausm_{self._suffix(Q)}", *map(_ptr, (Q, Fp, Fm, S, Ax, Ay, Az)))

    def _advect_comp(self, d, F, phi, S, Fp, Fm, NV):
        self._call(f"f3d_advect_{'xyz'[d]}_{self._suffix(Fp)}", *map(_ptr, (F, phi, S, Fp, Fm)), c_int32(NV))

    def advect_x(self, F, phi, S, Fp, Fm, NV=5):
        """schemes.jl:86-289 (component-wise)"""
        self._advect_comp(0, F, phi, S, Fp, Fm, NV)

    def advect_y(self, F, phi, S, Fp, Fm, NV=5):
        """schemes.jl:291-494 (component-wise)"""
        self._advect_comp(1, F, phi, S, Fp, Fm, NV)

    def advect_z(self, F, phi, S, Fp, Fm, NV=5):
        """schemes.jl:496-698 (component-wise)"""
        self._advect_comp(2, F, phi, S, Fp, Fm, NV)

    def _advect(self, d, F, phi, S, Fp, Fm, Q, Ax, Ay, Az):
        self._call(f"f3d_advect_{'xyz'[d]}c_{self._suffix(Q)}", *map(_ptr, (F, phi, S, Fp, Fm, Q, Ax, Ay, Az)))

    def advect_xc(self, F, phi, S, Fp, Fm, Q, Ax, Ay, Az):
        """schemes.jl:700-1012"""
        self._advect(0, F, phi, S, Fp, Fm, Q, Ax, Ay, Az)

    def advect_yc(self, F, phi, S, Fp, Fm, Q, Ax, Ay, Az):
        """schemes.jl:1014-1328"""
        self._advect(1, F, phi, S, Fp, Fm, Q, Ax, Ay, Az)

    def advect_zc(self, F, phi, S, Fp, Fm, Q, Ax, Ay, Az):
        """schemes.jl:1330-1643"""
        self._advect(2, F, phi, S, Fp, Fm, Q, Ax, Ay, Az)

    def viscousFlux(self, Fv_x, Fv_y, Fv_z, Q, dxidx, dxidy, dxidz, detadx, detady, detadz,
                    dzetadx, dzetady, dzetadz, J):
        """viscous.jl:1-134 (laminar, plus the eddy viscosity of params.les: viscous.jl:66-106)"""
        self._call(f"f3d_viscous_flux_{self._suffix(Q)}",
                   *map(_ptr, (Fv_x, Fv_y, Fv_z, Q, dxidx, dxidy, dxidz, detadx, detady, detadz,
                               dzetadx, dzetady, dzetadz, J)))

    def div(self, U, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, dt, J):
        """div.jl:1-50"""
        self._call(f"f3d_div_{self._suffix(Fx)}", *map(_ptr, (U, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z)),
                   c_double(dt), _ptr(J))

    def div_LTS(self, U, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, dt, J):
        """div.jl:52-101 (dt: per-cell float64 array of the cell shape)"""
        self._call(f"f3d_div_lts_{self._suffix(Fx)}", *map(_ptr, (U, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, dt, J)))

    def compute_dt(self, dt, Q, J, S1, S2, S3):
        """utils.jl:73-100 (dt: float64 array of the cell shape; interior cells written)"""
        self._call(f"f3d_compute_dt_{self._suffix(Q)}", *map(_ptr, (dt, Q, J, S1, S2, S3)))

    def _pre(self, d, Q, sc, rth):
        self._call(f"f3d_pre_{'xyz'[d]}_{self._suffix(Q)}", _ptr(sc), _ptr(Q), c_double(rth))

    def pre_x(self, Q, sc, rth):
        """utils.jl:102-124 (sc: float64 array of the cell shape)"""
        self._pre(0, Q, sc, rth)

    def pre_y(self, Q, sc, rth):
        """utils.jl:126-147"""
        self._pre(1, Q, sc, rth)

    def pre_z(self, Q, sc, rth):
        """utils.jl:149-169"""
        self._pre(2, Q, sc, rth)

    def filter_x(self, U, Un, sc, s0):
        """utils.jl:171-193"""
        self._call("f3d_filter_x", _ptr(U), _ptr(Un), _ptr(sc), c_double(s0))

    def filter_y(self, U, Un, sc, s0):
        """utils.jl:195-217"""
        self._call("f3d_filter_y", _ptr(U), _ptr(Un), _ptr(sc), c_double(s0))

    def filter_z(self, U, Un, sc, s0):
        """utils.jl:219-241"""
        self._call("f3d_filter_z", _ptr(U), _ptr(Un), _ptr(sc), c_double(s0))

    def linearFilter_x(self, U, Un, s0):
        """utils.jl:243-266"""
        self._call("f3d_linear_filter_x", _ptr(U), _ptr(Un), c_double(s0))

    def linearFilter_y(self, U, Un, s0):
        """utils.jl:268-291"""
        self._call("f3d_linear_filter_y", _ptr(U), _ptr(Un), c_double(s0))

    def linearFilter_z(self, U, Un, s0):
        """utils.jl:293-316"""
        self._call("f3d_linear_filter_z", _ptr(U), _ptr(Un), c_double(s0))

    def filter_step(self, U, Un, Q):
        """The filtering block of the time loop (solver.jl:446-470) with the params' settings."""
        prec = self._prec(Q)
        self._call(f"f3d_filter_step_{self._suffix(Q)}", self.workspace(prec).handle, _ptr(U), _ptr(Un), _ptr(Q))

    def linComb(self, U, Un, NV, a, b):
        """utils.jl:57-70"""
        self._call("f3d_lin_comb", _ptr(U), _ptr(Un), c_int32(NV), c_double(a), c_double(b))

    def c2Prim(self, U, Q):
        """utils.jl:2-30"""
        self._call(f"f3d_c2prim_{self._suffix(Q)}", _ptr(U), _ptr(Q))

    def prim2c(self, U, Q):
        """utils.jl:33-54"""
        self._call(f"f3d_prim2c_{self._suffix(Q)}", _ptr(U), _ptr(Q))

    def exchange_ghost(self, Q, NV):
        """mpi.jl:2-103 with one rank: local periodic copies (no MPI, no host staging)."""
        self._call(f"f3d_exchange_ghost_{self._suffix(Q)}", _ptr(Q), c_int32(NV))

    def fillGhost(self, Q, U, inflow=None):
        """boundary.jl:151-156 (fill_x, fill_y, fill_z) with the face kinds of params.bc."""
        self._call(f"f3d_fill_ghost_{self._suffix(Q)}", _ptr(Q), _ptr(U), _ptr(inflow))

    # -- finite-volume branch (FVM.jl) ---------------------------------------------------------
    def _fvm(self, name, Q, U, phi, S, F, x, y, z):
        self._call(f"f3d_{name}_{self._suffix(Q)}", *map(_ptr, (Q, U, phi, S, F, x, y, z)))

    def Eigen_reconstruct_i(self, Q, U, phi, S, Fx, x, y, z):
        """FVM.jl:6-238 (x, y, z: float64 mesh points)."""
        self._fvm("eigen_reconstruct_i", Q, U, phi, S, Fx, x, y, z)

    def Eigen_reconstruct_j(self, Q, U, phi, S, Fy, x, y, z):
        """FVM.jl:240-453"""
        self._fvm("eigen_reconstruct_j", Q, U, phi, S, Fy, x, y, z)

    def Eigen_reconstruct_k(self, Q, U, phi, S, Fz, x, y, z):
        """FVM.jl:455-670"""
        self._fvm("eigen_reconstruct_k", Q, U, phi, S, Fz, x, y, z)

    def Conser_reconstruct_i(self, Q, U, phi, S, Fx, x, y, z):
        """FVM.jl:672-865"""
        self._fvm("conser_reconstruct_i", Q, U, phi, S, Fx, x, y, z)

    def Conser_reconstruct_j(self, Q, U, phi, S, Fy, x, y, z):
        """FVM.jl:867-1044"""
        self._fvm("conser_reconstruct_j", Q, U, phi, S, Fy, x, y, z)

    def Conser_reconstruct_k(self, Q, U, phi, S, Fz, x, y, z):
        """FVM.jl:1046-1222"""
        self._fvm("conser_reconstruct_k", Q, U, phi, S, Fz, x, y, z)

    def flowAdvance(self, U, Q, Fp, Fm, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, s1, s2, s3, dxidx, dxidy, dxidz,
                    detadx, detady, detadz, dzetadx, dzetady, dzetadz, J, x, y, z, phi):
        """solver.jl:19: finite-volume branch (21-31) when params.fvm, else finite-difference (133-188);
        x, y, z are float64 mesh points in both numeric forms."""
        self._call(f"f3d_flow_advance_{self._suffix(Q)}",
                   *map(_ptr, (U, Q, Fp, Fm, Fx, Fy, Fz, Fv_x, Fv_y, Fv_z, s1, s2, s3, dxidx, dxidy, dxidz,
                               detadx, detady, detadz, dzetadx, dzetady, dzetadz, J, x, y, z, phi)))

    # -- fused path ---------------------------------------------------------------------------
    def _prec(self, Q):
        return PREC_F64 if self._suffix(Q) == "f64" else PREC_F32B

    def rhs(self, R, Q, metrics, J, phi=None):
        """R = J*(-dF + dFv) on interior cells (the div increment per unit dt)."""
        prec = self._prec(Q)
        ws = self.workspace(prec)
        fn = getattr(lib(), f"f3d_rhs_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, _ptr(R), _ptr(Q), _ptrs(metrics), _ptr(J), _ptr(phi),
                  self.stream))

    def rhs_noalloc(self, R, Q, metrics, J, phi=None):
        """The FP64 RHS without a workspace (f3d_rhs_noalloc_f64: the library's static per-GPU scratch)."""
        _check(lib().f3d_rhs_noalloc_f64(ctypes.byref(self.params), _ptr(R), _ptr(Q), _ptrs(metrics), _ptr(J),
                                         _ptr(phi), self.stream))

    def rk3_stage(self, KRK, U, Un, Q, metrics, J, dt, inflow=None):
        """One stage of the loop body solver.jl:416-444."""
        prec = self._prec(Q)
        ws = self.workspace(prec)
        fn = getattr(lib(), f"f3d_rk3_stage_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, c_int32(KRK), _ptr(U), _ptr(Un), _ptr(Q), _ptrs(metrics),
                  _ptr(J), c_double(dt), _ptr(inflow), self.stream))

    def time_step(self, nsteps, U, Un, Q, metrics, J, dt, inflow=None):
        """`nsteps` SSP-RK3 steps (solver.jl:410-470 without I/O): local time stepping and filtering per
        params (the filtering block after the steps of this call whose index is a multiple of
        filtering_interval)."""
        prec = self._prec(Q)
        ws = self.workspace(prec)
        fn = getattr(lib(), f"f3d_time_step_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, c_int32(nsteps), _ptr(U), _ptr(Un), _ptr(Q),
                  _ptrs(metrics), _ptr(J), c_double(dt), _ptr(inflow), self.stream))

    def fvm_geometry(self, x, y, z, geo):
        """Face geometry arrays geo[d] (params.face_shape(d) with 4 components: nx, ny, nz, Area) from
        the float64 mesh points, in the dtype of geo (float64 or float32)."""
        sfx = "f64" if geo[0].dtype == _torch().float64 else "f32b"
        self._call(f"f3d_fvm_geometry_{sfx}", _ptr(x), _ptr(y), _ptr(z), _ptrs(geo))

    def rhs_fvm(self, R, Q, U, metrics, J, geo, phi=None):
        """Finite-volume RHS (params.fvm = FVM_EIGEN / FVM_CONSER); U with ghosts, geo from fvm_geometry."""
        ws = self.workspace(self._prec(Q))
        fn = getattr(lib(), f"f3d_rhs_fvm_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, _ptr(R), _ptr(Q), _ptr(U), _ptrs(metrics), _ptr(J),
                  _ptrs(geo), _ptr(phi), self.stream))

    def rk3_stage_fvm(self, KRK, U, Un, Q, metrics, J, geo, dt, inflow=None):
        ws = self.workspace(self._prec(Q))
        fn = getattr(lib(), f"f3d_rk3_stage_fvm_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, c_int32(KRK), _ptr(U), _ptr(Un), _ptr(Q), _ptrs(metrics),
                  _ptr(J), _ptrs(geo), c_double(dt), _ptr(inflow), self.stream))

    def time_step_fvm(self, nsteps, U, Un, Q, metrics, J, geo, dt, inflow=None):
        ws = self.workspace(self._prec(Q))
        fn = getattr(lib(), f"f3d_time_step_fvm_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, c_int32(nsteps), _ptr(U), _ptr(Un), _ptr(Q),
                  _ptrs(metrics), _ptr(J), _ptrs(geo), c_double(dt), _ptr(inflow), self.stream))

    def cfl_dt(self, dt_out, Q, metrics, J, cfl):
        """Global CFL time step into the device scalar dt_out (float64 tensor of one element)."""
        prec = self._prec(Q)
        ws = self.workspace(prec)
        fn = getattr(lib(), f"f3d_cfl_dt_{self._suffix(Q)}")
        _check(fn(ctypes.byref(self.params), ws.handle, _ptr(dt_out), _ptr(Q), _ptrs(metrics), _ptr(J),
                  c_double(cfl), self.stream))

    def nan_check(self, U, NV=5):
        """any(isnan, U) (solver.jl:481-486): raises F3DNaNError when U holds NaN/Inf."""
        ws = self.workspace(PREC_F64) if PREC_F64 in self._ws else next(iter(self._ws.values()), None)
        if ws is None:
            ws = self.workspace(PREC_F64)
        _check(lib().f3d_nan_check_sync(ctypes.byref(self.params), ws.handle, _ptr(U), c_int32(NV), self.stream))

    def metrics(self, x, y, z, metrics, J, s=(None, None, None)):
        """FP64 metrics and Jacobian from node coordinates (parse_mesh.jl:139-248)."""
        _check(lib().f3d_metrics_f64(ctypes.byref(self.params), _ptr(x), _ptr(y), _ptr(z), _ptrs(metrics),
                                     _ptr(J), _ptrs(list(s)), self.stream))

    # -- adjoint (FP64, scheme mask frozen) ------------------------------------------------------
    def adjoint_workspace(self):
        if self._aws is None:
            self._aws = AdjointWorkspace(self.params)
        return self._aws

    def rhs_vjp(self, lamQ, lamR, Q, metrics, J):
        """lamQ (all cells, 6 vars) = (dR/dQ)^T lamR for R of `rhs`."""
        aws = self.adjoint_workspace()
        _check(lib().f3d_rhs_vjp_f64(ctypes.byref(self.params), aws.handle, _ptr(lamR), _ptr(Q), _ptrs(metrics),
                                     _ptr(J), _ptr(lamQ), self.stream))

    def ghost_vjp(self, lamQ, Q, lam_inflow=None, lam_bc=None):
        """Reverse of exchange_ghost + fillGhost: ghost adjoints folded into their sources."""
        self._call("f3d_ghost_vjp_f64", _ptr(lamQ), _ptr(Q), _ptr(lam_inflow), _ptr(lam_bc))

    def c2Prim_vjp(self, lamU, lamQ, U):
        self._call("f3d_c2prim_vjp_f64", _ptr(lamU), _ptr(lamQ), _ptr(U))

    def rk3_stage_vjp(self, KRK, lam_out, lam_s, lamUn, Us, Qs, metrics, J, dt, lam_inflow=None, lam_bc=None):
        """Reverse of `rk3_stage` KRK at its input state (Us, Qs)."""
        aws = self.adjoint_workspace()
        _check(lib().f3d_rk3_stage_vjp_f64(ctypes.byref(self.params), aws.handle, c_int32(KRK), _ptr(lam_out),
                                           _ptr(lam_s), _ptr(lamUn), _ptr(Us), _ptr(Qs), _ptrs(metrics), _ptr(J),
                                           c_double(dt), _ptr(lam_inflow), _ptr(lam_bc), self.stream))

    def adjoint(self, nsteps, every, U0, metrics, J, dt, W, obj, lamU0, inflow=None, lam_inflow=None, lam_bc=None,
                ckpt=None):
        """Checkpointed multi-step adjoint of sum_n dt <W, Q^{n+1}> (see f3d_adjoint_f64).
        `ckpt` is a float64 device tensor of at least adjoint_ckpt_bytes() bytes (allocated if None)."""
        import torch
        need = lib().f3d_adjoint_ckpt_bytes(ctypes.byref(self.params), nsteps, every)
        if ckpt is None:
            ckpt = torch.empty(need // 8, dtype=torch.float64, device=U0.device)
        aws = self.adjoint_workspace()
        _check(lib().f3d_adjoint_f64(ctypes.byref(self.params), aws.handle, c_int32(nsteps), c_int32(every),
                                     _ptr(U0), _ptrs(metrics), _ptr(J), c_double(dt), _ptr(inflow), _ptr(W),
                                     _ptr(ckpt), c_size_t(ckpt.numel() * ckpt.element_size()), _ptr(obj),
                                     _ptr(lamU0), _ptr(lam_inflow), _ptr(lam_bc), self.stream))
        return ckpt

    def adjoint_ckpt_bytes(self, nsteps, every):
        return lib().f3d_adjoint_ckpt_bytes(ctypes.byref(self.params), nsteps, every)


def _rhs_function():
    import torch

    class _RHS(torch.autograd.Function):
        """R = rhs(Q) (FP64) with its hand-written VJP as the backward (mask frozen)."""

        @staticmethod
        def forward(ctx, Q, f3d, metrics, J):
            R = torch.zeros(f3d.params.cell_shape(5), dtype=torch.float64, device=Q.device)
            f3d.rhs(R, Q, metrics, J)
            ctx.f3d, ctx.metrics = f3d, metrics
            ctx.save_for_backward(Q, J)
            return R

        @staticmethod
        def backward(ctx, gR):
            Q, J = ctx.saved_tensors
            lamQ = torch.empty_like(Q)
            ctx.f3d.rhs_vjp(lamQ, gR.contiguous(), Q, ctx.metrics, J)
            return lamQ, None, None, None

    return _RHS


class RHSFunction:
    """`RHSFunction.apply(Q, f3d, metrics, J)`: torch.autograd entry of the FP64 RHS."""

    @staticmethod
    def apply(Q, f3d, metrics, J):
        global _RHS_FN
        if _RHS_FN is None:
            _RHS_FN = _rhs_function()
        return _RHS_FN.apply(Q, f3d, metrics, J)


_RHS_FN = None
