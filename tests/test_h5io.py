"""The reference's on-disk formats through libf3d_h5.so (include/f3d_h5.h): mesh.h5 and metrics.h5 as
parse_mesh.jl:250-273 writes them and solver.jl:238-273 reads them, the chk-*.h5 checkpoint of
IO.jl:152-179 / solver.jl:211-222.

The layout is checked with HDF5's own h5dump (independent of this library): dataset names, HDF5 types,
dims in the reversed order of the Julia arrays, and single elements at known positions.  Round trips
are exact for Float64 files and equal the float32 rounding for Float32 files.  Ghost mesh points follow
the reference's linear extrapolation (solver.jl:682-829), exact on an affine mesh with dyadic spacing."""
import importlib
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

f3d = importlib.import_module("flame3d-mhd_amd")
h5 = importlib.import_module("flame3d-mhd_amd.h5io")
NG = 4
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
METRIC_NAMES = ["dξdx", "dξdy", "dξdz", "dηdx", "dηdy", "dηdz", "dζdx", "dζdy", "dζdz"]


def _dump(path, *args):
    r = subprocess.run([H5DUMP, *args, path], capture_output=True, text=True, check=True)
    return r.stdout


def _header(path):
    """{dataset name: (type, dims or 'SCALAR')} from h5dump -H."""
    out, cur = {}, None
    for line in _dump(path, "-H").splitlines():
        m = re.search(r'DATASET "(.+)"', line)
        if m:
            cur = m.group(1)
            continue
        m = re.search(r"DATATYPE\s+(\S+)", line)
        if m and cur:
            out[cur] = [m.group(1), None]
        m = re.search(r"DATASPACE\s+SIMPLE \{ \( ([0-9, ]+) \)", line)
        if m and cur:
            out[cur][1] = tuple(int(v) for v in m.group(1).split(","))
        if "DATASPACE  SCALAR" in line and cur:
            out[cur][1] = "SCALAR"
    return out


def _element(path, name, idx):
    out = _dump(path, "-d", f"/{name}", "-s", ",".join(map(str, idx)), "-c", ",".join("1" for _ in idx))
    m = re.search(r"\(([0-9,]+)\): ([-0-9.eE+]+)", out)
    return float(m.group(2))


def _affine_mesh(p):
    shp = p.cell_shape()
    k, j, i = np.meshgrid(*[np.arange(n, dtype=np.float64) - NG for n in shp], indexing="ij")
    return 0.125 * i + 0.0625 * j, 0.25 * j - 0.125 * k, 0.5 * k + 0.03125 * i


@pytest.fixture
def tmp(tmp_path):
    return str(tmp_path)


def test_mesh_layout_and_round_trip(tmp):
    p = f3d.make_params(6, 5, 5)
    x, y, z = _affine_mesh(p)
    path = os.path.join(tmp, "mesh.h5")
    h5.write_mesh(path, p, x, y, z)
    hdr = _header(path)
    for n in ("NG", "Nx", "Ny", "Nz"):
        assert hdr[n] == ["H5T_STD_I64LE", "SCALAR"], n
    assert hdr["coords"] == ["H5T_IEEE_F32LE", (5, 5, 6, 3)]  # Julia (3, Nx, Ny, Nz)
    assert h5.mesh_dims(path) == (4, 6, 5, 5)
    # Julia coords[c, i, j, k] (interior point i, j, k, 1-based) is file element (k-1, j-1, i-1, c-1)
    for (c, i, j, k) in ((1, 2, 1, 1), (2, 6, 5, 5), (3, 3, 4, 2)):
        ref = [x, y, z][c - 1][NG + k - 1, NG + j - 1, NG + i - 1]
        assert _element(path, "coords", (k - 1, j - 1, i - 1, c - 1)) == ref
    xr, yr, zr = h5.read_mesh(path, p)
    # dyadic affine coordinates: float32-exact, and the linear extrapolation of the ghosts is exact
    for a, b in ((xr, x), (yr, y), (zr, z)):
        assert np.array_equal(a, b)


def test_mesh_ghost_extrapolation_on_a_stretched_mesh(tmp):
    """Face ghosts are the point reflections 2 x_b - x_(2b - g) through the boundary point; xy corners
    x(i, jb) + x(ib, j) - x(ib, jb); z ghosts over the whole xy extent after those."""
    p = f3d.make_params(7, 6, 5)
    shp = p.cell_shape()
    k, j, i = np.meshgrid(*[np.arange(n, dtype=np.float64) - NG for n in shp], indexing="ij")
    x = i + 0.1 * np.sin(0.3 * j)
    y = np.sinh(0.4 * j) + 0.05 * i * i
    z = 0.7 * k + 0.01 * i * j
    path = os.path.join(tmp, "mesh.h5")
    h5.write_mesh(path, p, x, y, z, file_dtype=np.float64)
    got = h5.read_mesh(path, p)
    I = slice(NG, -NG)
    for a, src in zip(got, (x, y, z)):
        e = np.zeros_like(src)
        e[I, I, I] = src[I, I, I]
        nx, ny, nz = shp[2] - 2 * NG, shp[1] - 2 * NG, shp[0] - 2 * NG
        for g in range(NG):  # x faces: ghost g (0-based) mirrors 2NG - g through NG
            e[I, I, g] = 2 * e[I, I, NG] - e[I, I, 2 * NG - g]
            e[I, I, NG + nx + g] = 2 * e[I, I, NG + nx - 1] - e[I, I, NG + nx - 2 - g]
        for g in range(NG):
            e[I, g, I] = 2 * e[I, NG, I] - e[I, 2 * NG - g, I]
            e[I, NG + ny + g, I] = 2 * e[I, NG + ny - 1, I] - e[I, NG + ny - 2 - g, I]
        for xs, xb in ((slice(0, NG), NG), (slice(NG + nx, None), NG + nx - 1)):
            for ys, yb in ((slice(0, NG), NG), (slice(NG + ny, None), NG + ny - 1)):
                e[I, ys, xs] = e[I, yb:yb + 1, xs] + e[I, ys, xb:xb + 1] - e[I, yb:yb + 1, xb:xb + 1]
        for g in range(NG):
            e[g] = 2 * e[NG] - e[2 * NG - g]
            e[NG + nz + g] = 2 * e[NG + nz - 1] - e[NG + nz - 2 - g]
        assert np.array_equal(a, e)


@pytest.mark.parametrize("file_dtype", [np.float32, np.float64])
def test_metrics_layout_and_round_trip(tmp, file_dtype):
    p = f3d.make_params(6, 5, 4)
    rng = np.random.default_rng(1)
    m = [rng.standard_normal(p.cell_shape()) for _ in range(9)]
    J = 1 + rng.random(p.cell_shape())
    path = os.path.join(tmp, "metrics.h5")
    h5.write_metrics(path, p, m, J, file_dtype=file_dtype)
    hdr = _header(path)
    want = "H5T_IEEE_F32LE" if file_dtype == np.float32 else "H5T_IEEE_F64LE"
    for n in METRIC_NAMES + ["J"]:
        assert hdr[n] == [want, (12, 13, 14)], n  # Julia (Nx_tot, Ny_tot, Nz_tot)
    # element (i, j, k) 1-based of dηdz is file (k-1, j-1, i-1)
    assert _element(path, "dηdz", (3, 2, 7)) == pytest.approx(float(np.float32(m[5][3, 2, 7])) if file_dtype ==
                                                              np.float32 else m[5][3, 2, 7], rel=2e-6)  # h5dump prints ~7 digits
    mr, Jr = h5.read_metrics(path, p)
    rnd = (lambda a: a.astype(np.float32).astype(np.float64)) if file_dtype == np.float32 else (lambda a: a)
    for a, b in zip(mr + [Jr], m + [J]):
        assert np.array_equal(a, rnd(b))


@pytest.mark.parametrize("qdtype", [np.float64, np.float32])
@pytest.mark.parametrize("compress", [0, 1])
def test_checkpoint_layout_and_round_trip(tmp, qdtype, compress):
    p = f3d.make_params(6, 5, 4)
    Q = (1 + np.random.default_rng(2).random(p.cell_shape(6))).astype(qdtype)
    path = os.path.join(tmp, "chk-100.h5")
    h5.write_checkpoint(path, p, Q, compress=compress)
    hdr = _header(path)
    assert hdr["Q_h"] == ["H5T_IEEE_F32LE", (1, 6, 12, 13, 14)]  # Julia (Nx_tot, Ny_tot, Nz_tot, 6, 1)
    # Julia Q_h[i, j, k, n, 1] is file (0, n-1, k-1, j-1, i-1)
    assert _element(path, "Q_h", (0, 4, 2, 5, 9)) == pytest.approx(float(np.float32(Q[4, 2, 5, 9])), rel=2e-6)
    for rd in (np.float64, np.float32):
        Qr = h5.read_checkpoint(path, p, dtype=rd)
        assert np.array_equal(Qr, Q.astype(np.float32).astype(rd))


def test_errors_are_reported(tmp):
    p = f3d.make_params(6, 5, 5)
    x, y, z = _affine_mesh(p)
    path = os.path.join(tmp, "mesh.h5")
    h5.write_mesh(path, p, x, y, z)
    with pytest.raises(h5.F3DH5Error) as e:
        h5.read_mesh(path, f3d.make_params(6, 5, 6))
    assert e.value.code == -2 and "Nz" in str(e.value)
    small = f3d.make_params(6, 5, 4)
    h5.write_mesh(os.path.join(tmp, "small.h5"), small, *_affine_mesh(small))
    with pytest.raises(h5.F3DH5Error) as e:
        h5.read_mesh(os.path.join(tmp, "small.h5"), small)
    assert e.value.code == -2 and "NG+1" in str(e.value)
    with pytest.raises(h5.F3DH5Error) as e:
        h5.read_metrics(os.path.join(tmp, "missing.h5"), p)
    assert e.value.code == -7
    with pytest.raises(h5.F3DH5Error) as e:
        h5.read_metrics(path, p)  # a mesh file has no metrics
    assert e.value.code == -7 and "dξdx" in str(e.value)
    Q = np.ones(p.cell_shape(6))
    cpath = os.path.join(tmp, "chk.h5")
    h5.write_checkpoint(cpath, p, Q)
    with pytest.raises(h5.F3DH5Error) as e:
        h5.read_checkpoint(cpath, f3d.make_params(6, 6, 4))
    assert e.value.code == -2


def test_library_exports_every_declared_symbol():
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "f3d_h5.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(f3d_h5_[a-z0-9_]+)\s*\(", src)))
    assert len(names) == 8
    out = subprocess.run(["nm", "-D", "--defined-only", h5.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s(f3d_h5_\w+)", out.stdout))
    assert not [n for n in names if n not in exported]
