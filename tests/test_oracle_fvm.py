"""Pins of the CPU oracle's finite-volume branch (FVM.jl) by closed forms and invariants.

Nothing here re-evaluates the oracle's own formulas: the face geometry is checked against the
metric cofactors (exact on affine meshes, second-order convergent on a curved one); the HLLC flux
against the Euler flux, its mirror and rotation symmetries, exact contact/shear resolution, and
the algebraically different p*-form of the star-region flux (Batten et al. 1997); the
reconstructions against the exact face value of polynomial cell averages.
"""
import importlib

import numpy as np
import pytest

from oracle_lib import Oracle

f3d = importlib.import_module("flame3d-mhd_amd")
NG = 4
GAMMA = 1.4


def params(N=(8, 7, 6), **kw):
    return f3d.make_params(*N, **kw)


def mesh_points(p, fn):
    """x, y, z at every mesh index (0-based, ghosts included) from fn(i, j, k) -> (x, y, z)."""
    k, j, i = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in (p.Nzp + 8, p.Nyp + 8, p.Nxp + 8)),
                          indexing="ij")
    return [np.ascontiguousarray(a) for a in fn(i, j, k)]


def face_shape(p, d, nv):
    n = [p.Nxp, p.Nyp, p.Nzp]
    n[d] += 1
    return (nv, n[2], n[1], n[0])


def geometry(p, d, X):
    geo = np.zeros(face_shape(p, d, 4))
    Oracle(p).fvm_geometry(d, *X, geo)
    return geo


# ----------------------------------------------------------------------------- geometry
def test_geometry_affine_equals_metric_rows():
    """x = M (i, j, k) + b: n*Area of every face equals the J^-1-scaled metric row of its direction
    (the cofactor column of M), for i, j and k faces alike (pins the j-face orientation flip)."""
    p = params()
    rng = np.random.default_rng(3)
    M = np.eye(3) * 0.7 + 0.2 * rng.standard_normal((3, 3))
    assert np.linalg.det(M) > 0.1
    b = np.array([3.0, -2.0, 5.0])
    X = mesh_points(p, lambda i, j, k: [M[c, 0] * i + M[c, 1] * j + M[c, 2] * k + b[c] for c in range(3)])
    cof = np.linalg.det(M) * np.linalg.inv(M)  # rows: grad xi_d / J
    m = [np.zeros(p.cell_shape()) for _ in range(9)]
    J = np.zeros(p.cell_shape())
    Oracle(p).metrics(*X, m, J)
    for d in range(3):
        geo = geometry(p, d, X)
        nA = geo[:3] * geo[3]
        for c in range(3):
            assert np.allclose(nA[c], cof[d, c], rtol=0, atol=1e-12 * np.abs(cof).max())
            assert np.allclose(m[3 * d + c][NG:-NG, NG:-NG, NG:-NG], cof[d, c], atol=1e-12 * np.abs(cof).max())
        assert np.allclose(np.sum(geo[:3] ** 2, axis=0), 1.0, atol=1e-14)


def _curved(i, j, k, h):
    s, t, u = i * h, j * h, k * h
    x = s + 0.1 * np.sin(t + 0.3) * np.cos(u)
    y = t + 0.15 * np.sin(s) + 0.05 * u * u
    z = u + 0.1 * np.sin(s + t)
    return x, y, z


def _curved_jac(s, t, u):
    """Columns dX/ds, dX/dt, dX/du of _curved (h = 1 units of the physical parameter)."""
    xs, xt, xu = np.ones_like(s), 0.1 * np.cos(t + 0.3) * np.cos(u), -0.1 * np.sin(t + 0.3) * np.sin(u)
    ys, yt, yu = 0.15 * np.cos(s), np.ones_like(s), 0.1 * u
    zs, zt, zu = 0.1 * np.cos(s + t), 0.1 * np.cos(s + t), np.ones_like(s)
    return np.array([[xs, xt, xu], [ys, yt, yu], [zs, zt, zu]])


@pytest.mark.parametrize("d", [0, 1, 2])
def test_geometry_second_order_on_curved_mesh(d):
    """n*Area -> (x_a x x_b) h^2 at the face centre (the cofactor of the mapping), error O(h^2)."""
    errs = []
    for N in (8, 16):
        p = params((N, N, N))
        h = 1.0 / N
        X = mesh_points(p, lambda i, j, k: _curved(i - NG, j - NG, k - NG, h))
        geo = geometry(p, d, X)
        nA = geo[:3] * geo[3]
        # face f (0-based in its array) sits between cells NG+f-1 and NG+f along d
        shp = face_shape(p, d, 1)[1:]
        kk, jj, ii = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shp), indexing="ij")
        idx = [ii, jj, kk]
        idx[d] = idx[d] - 0.5
        s, t, u = (a * h for a in idx)
        Jm = _curved_jac(s, t, u)
        a, b = [(1, 2), (2, 0), (0, 1)][d]
        exact = np.cross(Jm[:, a].transpose(1, 2, 3, 0), Jm[:, b].transpose(1, 2, 3, 0)).transpose(3, 0, 1, 2) * h * h
        errs.append(np.abs(nA - exact).max() / (h * h))
    assert errs[1] < errs[0] / 3.2, errs


# ----------------------------------------------------------------------------- HLLC
def euler_flux(U, n):
    rho, mx, my, mz, E = U
    u, v, w = mx / rho, my / rho, mz / rho
    pr = (GAMMA - 1) * (E - 0.5 * rho * (u * u + v * v + w * w))
    q = u * n[0] + v * n[1] + w * n[2]
    return np.array([rho * q, mx * q + pr * n[0], my * q + pr * n[1], mz * q + pr * n[2], (E + pr) * q])


def cons(rho, u, v, w, pr):
    return np.array([rho, rho * u, rho * v, rho * w, pr / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def rand_states(rng, nf, mach=0.6):
    rho = 0.5 + rng.random(nf)
    pr = 0.5 + rng.random(nf)
    c = np.sqrt(GAMMA * pr / rho)
    vel = mach * c * (2 * rng.random((3, nf)) - 1)
    return cons(rho, *vel, pr)


def unit(rng, nf):
    n = rng.standard_normal((3, nf))
    return n / np.linalg.norm(n, axis=0)


def hllc(p, UL, UR, n):
    F = np.zeros_like(UL)
    Oracle(p).hllc(np.ascontiguousarray(UL), np.ascontiguousarray(UR), np.ascontiguousarray(n), F)
    return F


def test_hllc_consistency():
    p = params()
    rng = np.random.default_rng(5)
    for mach in (0.3, 0.9, 2.5):
        U = rand_states(rng, 200, mach)
        n = unit(rng, 200)
        F = hllc(p, U, U, n)
        assert np.allclose(F, euler_flux(U, n), rtol=0, atol=1e-10 * np.abs(euler_flux(U, n)).max())


def test_hllc_mirror_and_rotation():
    p = params()
    rng = np.random.default_rng(6)
    UL, UR = rand_states(rng, 300, 1.2), rand_states(rng, 300, 1.2)
    n = unit(rng, 300)
    F = hllc(p, UL, UR, n)
    scale = np.abs(F).max()
    assert np.allclose(hllc(p, UR, UL, -n), -F, atol=1e-13 * scale)
    th = 0.7
    Rm = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]]) @ \
        np.array([[1, 0, 0], [0, np.cos(1.1), -np.sin(1.1)], [0, np.sin(1.1), np.cos(1.1)]])

    def rot(U):
        V = U.copy()
        V[1:4] = Rm @ U[1:4]
        return V
    Fr = hllc(p, rot(UL), rot(UR), Rm @ n)
    assert np.allclose(Fr, rot(F), atol=1e-12 * scale)


def test_hllc_contact_and_shear_exact():
    """Equal pressure and normal velocity, jumps in density and tangential velocity: the exact
    flux is the upwind state's Euler flux (HLLC resolves contact and shear waves)."""
    p = params()
    rng = np.random.default_rng(7)
    nf = 200
    n = unit(rng, nf)
    pr = 0.5 + rng.random(nf)
    q = 0.8 * (2 * rng.random(nf) - 1)
    tL, tR = rng.standard_normal((3, nf)), rng.standard_normal((3, nf))
    tL -= n * np.sum(tL * n, axis=0)
    tR -= n * np.sum(tR * n, axis=0)
    vL, vR = q * n + 0.3 * tL, q * n + 0.3 * tR
    UL = cons(0.5 + rng.random(nf), *vL, pr)
    UR = cons(0.5 + rng.random(nf), *vR, pr)
    F = hllc(p, UL, UR, n)
    up = np.where(q >= 0, euler_flux(UL, n), euler_flux(UR, n))
    assert np.allclose(F, up, atol=1e-10 * np.abs(up).max())


def _batten_hllc(UL, UR, n):
    """Star-region flux in the p*-form (Batten, Clarke, Lambert, Causon 1997):
    F*_K = (S* (S_K U_K - F_K) + S_K p* D*) / (S_K - S*), p* = p_K + rho_K (S_K - q_K)(S* - q_K),
    D* = (0, n, S*), with the reference's Roe-averaged wave speeds and S* (FVM.jl:1250-1280)."""
    def prim(U):
        rho = U[0]
        u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
        pr = (GAMMA - 1) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
        return rho, u, v, w, pr
    rL, uL, vL, wL, pL = prim(UL)
    rR, uR, vR, wR, pR = prim(UR)
    qL = uL * n[0] + vL * n[1] + wL * n[2]
    qR = uR * n[0] + vR * n[1] + wR * n[2]
    cL, cR = np.sqrt(GAMMA * pL / rL), np.sqrt(GAMMA * pR / rR)
    a, b = np.sqrt(rL), np.sqrt(rR)
    ur, vr, wr = ((a * uL + b * uR) / (a + b), (a * vL + b * vR) / (a + b), (a * wL + b * wR) / (a + b))
    Hr = (a * (UL[4] + pL) / rL + b * (UR[4] + pR) / rR) / (a + b)
    cr = np.sqrt((GAMMA - 1) * (Hr - 0.5 * (ur * ur + vr * vr + wr * wr)))
    qr = ur * n[0] + vr * n[1] + wr * n[2]
    SL, SR = np.minimum(qL - cL, qr - cr), np.maximum(qR + cR, qr + cr)
    Ss = (pR - pL + rL * qL * (SL - qL) - rR * qR * (SR - qR)) / (rL * (SL - qL) - rR * (SR - qR) + 1e-12)
    FL, FR = euler_flux(UL, n), euler_flux(UR, n)
    Dst = np.array([np.zeros_like(Ss), n[0], n[1], n[2], Ss])
    pstL = pL + rL * (SL - qL) * (Ss - qL)
    pstR = pR + rR * (SR - qR) * (Ss - qR)
    FsL = (Ss * (SL * UL - FL) + SL * pstL * Dst) / (SL - Ss)
    FsR = (Ss * (SR * UR - FR) + SR * pstR * Dst) / (SR - Ss)
    return np.where(SL >= 0, FL, np.where(SR <= 0, FR, np.where(Ss >= 0, FsL, FsR)))


def test_hllc_star_states_against_p_star_form():
    p = params()
    rng = np.random.default_rng(8)
    nf = 400
    UL, UR = rand_states(rng, nf, 0.8), rand_states(rng, nf, 0.8)
    n = unit(rng, nf)
    F = hllc(p, UL, UR, n)
    ref = _batten_hllc(UL, UR, n)
    assert np.allclose(F, ref, atol=1e-11 * np.abs(ref).max())


# ----------------------------------------------------------------------------- reconstruction
def _cell_avg_poly(coef, xc):
    """Cell averages over [xc - 1/2, xc + 1/2] of sum_k coef[k] x^k."""
    out = np.zeros_like(xc)
    for kk, a in enumerate(coef):
        out += a * ((xc + 0.5) ** (kk + 1) - (xc - 0.5) ** (kk + 1)) / (kk + 1)
    return out


def _poly(coef, x):
    return sum(a * x ** kk for kk, a in enumerate(coef))


SCHEMES = [("up7", 0.0, 6), ("weno7", 0.5, 3), ("weno5", 5.0, 2), ("minmod", 50.0, 1)]


def _recon_case(p, d, deg, seed):
    """U = cell averages of degree-`deg` polynomials along d (scaled to keep rho, p > 0), Cartesian
    unit mesh; returns U, Q, S and the exact face values per variable as a function of the face
    coordinate."""
    rng = np.random.default_rng(seed)
    shp = p.cell_shape()
    idx = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shp), indexing="ij")  # k, j, i
    xd = idx[2 - d] - NG  # cell centre coordinate along d (cell NG at 0)
    L = float(shp[2 - d])
    base = np.array([1.0, 0.3, -0.2, 0.1, 2.5])
    coefs = []
    U = np.zeros(p.cell_shape(5))
    for m in range(5):
        c = np.zeros(deg + 1)
        c[0] = base[m]
        for kk in range(1, deg + 1):
            c[kk] = 0.05 * (2 * rng.random() - 1) / L ** kk
        coefs.append(c)
        U[m] = _cell_avg_poly(c, xd)
    Q = np.zeros(p.cell_shape(6))
    rho = U[0]
    u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
    pr = (GAMMA - 1) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    Q[:] = [rho, u, v, w, pr, pr / (rho * p.Rg)]
    S = np.ones(p.cell_shape())
    return U, Q, S, coefs


@pytest.mark.parametrize("eigen", [True, False])
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("name,phi0,deg", SCHEMES)
def test_reconstruction_exact_on_polynomials(eigen, d, name, phi0, deg):
    p = params((9, 8, 7))
    U, Q, S, coefs = _recon_case(p, d, deg, seed=11 + d)
    phi = np.full(p.cell_shape(), phi0)
    X = mesh_points(p, lambda i, j, k: [i - NG, j - NG, k - NG])
    F = np.zeros(face_shape(p, d, 5))
    UL, UR, geo = np.zeros_like(F), np.zeros_like(F), np.zeros(face_shape(p, d, 4))
    Oracle(p).fvm_reconstruct(d, eigen, Q, U, phi, S, F, *X, UL, UR, geo)
    # face f (0-based) between cells NG+f-1 and NG+f: coordinate f - 1/2 along d
    nfd = [p.Nxp, p.Nyp, p.Nzp][d] + 1
    xf = np.arange(nfd, dtype=np.float64) - 0.5
    sh = [1, 1, 1]
    sh[2 - d] = nfd
    exact = np.stack([np.broadcast_to(_poly(c, xf).reshape(sh), F.shape[1:]) for c in coefs])
    tol = 1e-12 if name == "up7" else 1e-10
    assert np.abs(UL - exact).max() < tol * np.abs(exact).max(), name
    if name == "minmod":
        # right state as written in the reference: first order (the right cell's value) on monotone data
        right = np.stack([np.broadcast_to(_cell_avg_poly(c, xf + 0.5).reshape(sh), F.shape[1:]) for c in coefs])
        assert np.abs(UR - right).max() < 1e-12 * np.abs(right).max()
    else:
        assert np.abs(UR - exact).max() < tol * np.abs(exact).max(), name
    # unit Cartesian mesh: n = e_d, Area = 1; F = Area * HLLC(UL, UR, n)
    assert np.allclose(geo[3], 1.0) and np.allclose(geo[d], 1.0)
    if name != "minmod":
        n = np.zeros((3,) + F.shape[1:])
        n[d] = 1.0
        Fe = euler_flux(exact, n)
        assert np.abs(F - Fe).max() < 1e-9 * np.abs(Fe).max()


def test_scheme_selection_uses_six_cell_sensor_window():
    """A single sensor spike at cell c switches exactly the faces whose window c-2..c+3 of the
    left cell contains it (FVM.jl:34): faces with left cell in [c-3, c+2]."""
    p = params((14, 6, 5))
    U, Q, S, coefs = _recon_case(p, 0, 4, seed=3)  # degree 4: UP7 exact, minmod not
    phi = np.zeros(p.cell_shape())
    c = NG + 7
    phi[:, :, c] = 50.0
    X = mesh_points(p, lambda i, j, k: [i - NG, j - NG, k - NG])
    F = np.zeros(face_shape(p, 0, 5))
    UL = np.zeros_like(F)
    Oracle(p).fvm_reconstruct(0, False, Q, U, phi, S, F, *X, UL, np.zeros_like(F), None)
    xf = np.arange(p.Nxp + 1, dtype=np.float64) - 0.5
    ex = _poly(coefs[0], xf)
    err = np.abs(UL[0, 2, 2, :] - ex)
    left_cell = np.arange(p.Nxp + 1) + NG - 1
    hit = (left_cell >= c - 3) & (left_cell <= c + 2)
    assert np.all(err[~hit] < 1e-12) and np.all(err[hit] > 1e-9)


def _fvm_case(p, X, U0, seed=0, noise=0.0):
    rng = np.random.default_rng(seed)
    U = np.zeros(p.cell_shape(5))
    for m in range(5):
        U[m] = U0[m] * (1 + noise * (2 * rng.random(p.cell_shape()) - 1))
    Q = np.zeros(p.cell_shape(6))
    rho = U[0]
    u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
    pr = (GAMMA - 1) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))
    Q[:] = [rho, u, v, w, pr, pr / (rho * p.Rg)]
    m = [np.zeros(p.cell_shape()) for _ in range(9)]
    J = np.zeros(p.cell_shape())
    Oracle(p).metrics(*X, m, J)
    return U, Q, m, J


@pytest.mark.parametrize("fvm", [1, 2])
def test_free_stream_preserved_on_affine_mesh(fvm):
    """Uniform state on a sheared affine mesh: the face area vectors of each cell close exactly,
    so the inviscid finite-volume RHS vanishes (to rounding); the viscous flux of a uniform state
    is zero."""
    p = params((8, 7, 6), fvm=fvm, bc=[0] * 6)
    M = np.array([[0.9, 0.2, 0.1], [-0.1, 1.1, 0.15], [0.05, -0.2, 0.8]])
    X = mesh_points(p, lambda i, j, k: [M[c, 0] * i + M[c, 1] * j + M[c, 2] * k for c in range(3)])
    U, Q, m, J = _fvm_case(p, X, cons(1.2, 0.4, -0.3, 0.2, 1.0))
    R = np.zeros(p.cell_shape(5))
    Oracle(p).rhs_fvm(R, Q, U, m, J, X)
    A = np.zeros_like(R)
    Oracle(p).rhs_absterms_fvm(A, Q, U, m, J, X)
    assert np.abs(R).max() < 1e-13 * A.max()


@pytest.mark.parametrize("fvm", [1, 2])
def test_periodic_conservation(fvm):
    """All-periodic box with a perturbed state: sum over interior cells of R/J telescopes to zero
    for every conserved variable (inviscid and viscous fluxes are single-valued on faces)."""
    p = params((10, 8, 6), fvm=fvm, bc=[0] * 6, viscous=1)
    X = mesh_points(p, lambda i, j, k: [0.1 * i, 0.12 * j, 0.09 * k])
    U, Q, m, J = _fvm_case(p, X, cons(1.0, 0.3, 0.1, -0.2, 2.0), seed=4, noise=0.2)
    o = Oracle(p)
    o.exchange(Q, 6)
    o.fill(Q, U, None)  # periodic ghosts of Q and U
    R = np.zeros(p.cell_shape(5))
    o.rhs_fvm(R, Q, U, m, J, X)
    A = np.zeros_like(R)
    o.rhs_absterms_fvm(A, Q, U, m, J, X)
    s = (slice(NG, -NG),) * 3
    for n in range(5):
        tot = np.sum(R[n][s] / J[s])
        assert abs(tot) < 1e-12 * np.sum(A[n][s] / J[s]), (n, tot)


# ----------------------------------------------------------------------------- float form
def test_difference_form_indicators_equal_reference_forms():
    """The float form's WENO7 indicators (quadratic forms of first differences) are the reference's
    quadratic forms in V: identical on integer stencils, where double arithmetic is exact."""
    import ctypes
    from oracle_lib import olib
    rng = np.random.default_rng(12)
    L = olib()
    for _ in range(200):
        V = rng.integers(-300, 300, 7).astype(np.float64)
        a, b = np.zeros(4), np.zeros(4)
        L.oracle_weno7_is_f64(0, ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(a.ctypes.data))
        L.oracle_weno7_is_f64(1, ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(b.ctypes.data))
        assert np.array_equal(a, b), (V, a, b)


@pytest.mark.parametrize("eigen", [True, False])
def test_float_form_states_track_fp64(eigen):
    """Reconstructed states of the float form on the wall-normal faces of the swept ramp (WENO7
    faces on smooth boundary-layer profiles, E ~ 1e4 next to rho ~ 0.05) stay near the FP64 ones.
    With the indicators evaluated in V the weights follow rounding noise there: the conser density
    and energy states were off by 3e-5 and 5e-5, the eigen transverse momenta by 3e-3."""
    from casegen import ramp_case
    c = ramp_case(20, 12, 10, sweep=30.0, dtype=np.float32, jump=0.5, fvm=1 if eigen else 2)
    p, o = c.params, Oracle(c.params)
    U = c.U()
    phi = np.zeros(p.cell_shape(), np.float32)
    o.shock_sensor(phi, c.Q)
    Q64 = c.Q.astype(np.float64)
    phi64 = np.zeros(p.cell_shape())
    o.shock_sensor(phi64, Q64)
    F32 = np.zeros(p.face_shape(1), np.float32)
    L32 = np.zeros_like(F32)
    o.fvm_reconstruct(1, eigen, c.Q, U, phi, c.s[1], F32, *c.coords, L32, np.zeros_like(F32))
    F64 = np.zeros(p.face_shape(1))
    L64 = np.zeros_like(F64)
    o.fvm_reconstruct(1, eigen, Q64, U, phi64, c.s[1].astype(np.float64), F64, *c.coords, L64, np.zeros_like(F64))
    rel = [np.abs(L32[n] - L64[n]).max() / max(np.abs(L64[n]).max(), 1e-30) for n in range(5)]
    for n in (0, 1, 4):
        assert rel[n] < 1e-6 if not eigen else rel[n] < 3e-6, (n, rel)
    if eigen:
        assert max(rel[2], rel[3]) < 1e-4, rel
