"""The reference's on-disk formats (mesh.h5, metrics.h5, chk-*.h5) through libf3d_h5.so (include/f3d_h5.h).

Arrays are numpy, C-contiguous, in the reversed (Julia column-major) shape used everywhere in the package:
a Julia (Nx_tot, Ny_tot, Nz_tot) array is a (Nz_tot, Ny_tot, Nx_tot) array here, Q is (6, Nz_tot, Ny_tot,
Nx_tot).  Move them to the GPU with torch.from_numpy(...).to("cuda:0").

  read_mesh(path)             -> params-shaped x, y, z with ghosts (run.jl:79-82, solver.jl:253-273, 682-829)
  write_mesh(path, p, x, y, z)                    (parse_mesh.jl:262-273)
  read_metrics(path, p)       -> 9 metrics, J     (solver.jl:238-251)
  write_metrics(path, p, metrics, J)              (parse_mesh.jl:250-261)
  read_checkpoint(path, p, dtype) -> Q            (solver.jl:211-222)
  write_checkpoint(path, p, Q)                    (IO.jl:152-179)
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libf3d_h5.so")
PREC_F64, PREC_F32B = 0, 1
_lib = None


class F3DH5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"f3d_h5 error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run flame3d-mhd_amd/build_lib.py (needs the HDF5 headers)")
        L = ctypes.CDLL(LIB_PATH)
        L.f3d_h5_last_error.restype = c_char_p
        L.f3d_h5_mesh_dims.argtypes = [c_char_p, POINTER(c_int64)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise F3DH5Error(rc, lib().f3d_h5_last_error().decode(errors="replace"))


def _p(a):
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return c_void_p(a.ctypes.data)


def _pp(arrs):
    return (c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _prec(dtype):
    return PREC_F64 if np.dtype(dtype) == np.float64 else PREC_F32B


def mesh_dims(path):
    """(NG, Nx, Ny, Nz) of a mesh.h5."""
    d = (c_int64 * 4)()
    _check(lib().f3d_h5_mesh_dims(path.encode(), d))
    return tuple(int(v) for v in d)


def read_mesh(path, params):
    """Mesh points with ghosts (float64, cell-shaped) for params matching the file."""
    shp = params.cell_shape()
    x, y, z = (np.zeros(shp) for _ in range(3))
    _check(lib().f3d_h5_read_mesh(path.encode(), ctypes.byref(params), _p(x), _p(y), _p(z)))
    return x, y, z


def write_mesh(path, params, x, y, z, file_dtype=np.float32, compress=3):
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    _check(lib().f3d_h5_write_mesh(path.encode(), ctypes.byref(params), _p(x), _p(y), _p(z),
                                   c_int32(_prec(file_dtype)), c_int32(compress)))


def read_metrics(path, params):
    """The 9 metrics in the reference's order and J, float64 with ghosts."""
    shp = params.cell_shape()
    m = [np.zeros(shp) for _ in range(9)]
    J = np.zeros(shp)
    _check(lib().f3d_h5_read_metrics(path.encode(), ctypes.byref(params), _pp(m), _p(J)))
    return m, J


def write_metrics(path, params, metrics, J, file_dtype=np.float32, compress=3):
    m = [np.ascontiguousarray(a, dtype=np.float64) for a in metrics]
    J = np.ascontiguousarray(J, dtype=np.float64)
    _check(lib().f3d_h5_write_metrics(path.encode(), ctypes.byref(params), _pp(m), _p(J),
                                      c_int32(_prec(file_dtype)), c_int32(compress)))


def read_checkpoint(path, params, dtype=np.float64):
    """Q (6 variables with ghosts) of rank 1 as `dtype` (float64 or float32)."""
    Q = np.zeros(params.cell_shape(6), dtype=dtype)
    _check(lib().f3d_h5_read_checkpoint(path.encode(), ctypes.byref(params), _p(Q), c_int32(_prec(dtype))))
    return Q


def write_checkpoint(path, params, Q, compress=1):
    Q = np.ascontiguousarray(Q)
    assert Q.dtype in (np.float64, np.float32) and Q.shape == params.cell_shape(6)
    _check(lib().f3d_h5_write_checkpoint(path.encode(), ctypes.byref(params), _p(Q), c_int32(_prec(Q.dtype)),
                                         c_int32(compress)))
