"""Pins of the oracle's other flux-vector splittings (split.jl:70-255) and of the component-wise
advect_x/y/z (schemes.jl:86-698) by closed forms: consistency Fp + Fm = the Euler flux on the metric
row, one-sided fluxes for |M| >= 1, continuity of the subsonic formulas at |M| = 1, the upwind signs
of the split mass fluxes, the Lax-Friedrichs dissipation, and exact reconstruction of polynomial cell
averages by each scheme of the component-wise advect."""
import importlib

import numpy as np
import pytest

from oracle_lib import Oracle

f3d = importlib.import_module("flame3d-mhd_amd")
NG = 4
GAMMA = 1.4


def _state(p, rng, mach_lo, mach_hi):
    """Random cells: rho, p > 0, velocity of Mach number in [mach_lo, mach_hi] along a random metric
    row A (|A| = S), random transverse components."""
    shp = p.cell_shape()
    rho = 0.5 + rng.random(shp)
    pr = 0.5 + rng.random(shp)
    c = np.sqrt(GAMMA * pr / rho)
    A = rng.standard_normal((3,) + shp)
    S = np.sqrt(np.sum(A * A, axis=0))
    nhat = A / S
    M = mach_lo + (mach_hi - mach_lo) * rng.random(shp)
    t = rng.standard_normal((3,) + shp)
    t -= nhat * np.sum(t * nhat, axis=0)
    vel = M * c * nhat + 0.3 * c * t
    Q = np.zeros(p.cell_shape(6))
    Q[0], Q[1:4], Q[4], Q[5] = rho, vel, pr, pr / (rho * p.Rg)
    return Q, A, S, M


def _euler(Q, A):
    rho, u, v, w, pr = Q[0], Q[1], Q[2], Q[3], Q[4]
    un = A[0] * u + A[1] * v + A[2] * w
    E = pr / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.array([rho * un, rho * u * un + A[0] * pr, rho * v * un + A[1] * pr, rho * w * un + A[2] * pr,
                     (E + pr) * un]), np.array([rho, rho * u, rho * v, rho * w, E]), un


def _split(p, method, Q, A, S):
    Fp = np.zeros(p.cell_shape(5))
    Fm = np.zeros_like(Fp)
    Oracle(p).flux_split_method(method, np.ascontiguousarray(Q), Fp, Fm, np.ascontiguousarray(S),
                                *[np.ascontiguousarray(a) for a in A])
    return Fp, Fm


@pytest.mark.parametrize("method", [f3d.SPLIT_SW, f3d.SPLIT_LF, f3d.SPLIT_VL, f3d.SPLIT_AUSM])
def test_split_consistency(method):
    p = f3d.make_params(6, 5, 4)
    rng = np.random.default_rng(3 + method)
    for lo, hi in ((-0.95, 0.95), (-3.0, 3.0)):
        Q, A, S, _ = _state(p, rng, lo, hi)
        Fp, Fm = _split(p, method, Q, A, S)
        F, _, _ = _euler(Q, A)
        for n in range(5):
            assert np.abs(Fp[n] + Fm[n] - F[n]).max() < 1e-12 * np.abs(F[n]).max(), (method, n)


@pytest.mark.parametrize("method", [f3d.SPLIT_VL, f3d.SPLIT_AUSM])
def test_upwind_sides_and_continuity(method):
    p = f3d.make_params(6, 5, 4)
    rng = np.random.default_rng(11)
    Q, A, S, M = _state(p, rng, 1.0, 2.0)
    Fp, Fm = _split(p, method, Q, A, S)
    F, _, _ = _euler(Q, A)
    assert np.all(Fm == 0) and np.allclose(Fp, F, rtol=0, atol=1e-14 * np.abs(F).max())
    Q, A, S, M = _state(p, rng, -2.0, -1.0)
    Fp, Fm = _split(p, method, Q, A, S)
    assert np.all(Fp == 0)
    # subsonic: split mass fluxes carry the upwind signs
    Q, A, S, M = _state(p, rng, -0.99, 0.99)
    Fp, Fm = _split(p, method, Q, A, S)
    assert np.all(Fp[0] >= 0) and np.all(Fm[0] <= 0)
    # the subsonic formulas meet the one-sided fluxes at M = +-1: a state just below M = 1 splits
    # into Fp within O(1e-6) of the full flux and Fm within O(1e-6) of zero
    for target in (1.0, -1.0):
        Q, A, S, M = _state(p, rng, target - 1e-7 * np.sign(target) * 2, target - 1e-7 * np.sign(target))
        Fp, Fm = _split(p, method, Q, A, S)
        F, _, _ = _euler(Q, A)
        up, down = (Fp, Fm) if target > 0 else (Fm, Fp)
        scale = np.abs(F).max()
        assert np.abs(up - F).max() < 1e-5 * scale and np.abs(down).max() < 1e-5 * scale, target


def test_lax_friedrichs_dissipation():
    p = f3d.make_params(6, 5, 4)
    Q, A, S, M = _state(p, np.random.default_rng(5), -2.0, 2.0)
    Fp, Fm = _split(p, f3d.SPLIT_LF, Q, A, S)
    F, U, un = _euler(Q, A)
    lam = np.abs(un) + np.sqrt(GAMMA * Q[4] / Q[0]) * S
    for n in range(5):
        assert np.abs((Fp[n] - Fm[n]) - lam * U[n]).max() < 1e-12 * np.abs(lam * U[n]).max()


def _cell_avg_poly(coef, xc):
    out = np.zeros_like(xc)
    for kk, a in enumerate(coef):
        out += a * ((xc + 0.5) ** (kk + 1) - (xc - 0.5) ** (kk + 1)) / (kk + 1)
    return out


@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("phi0,deg", [(0.0, 6), (0.5, 3), (5.0, 2), (50.0, 1)])
def test_component_advect_exact_on_polynomials(d, phi0, deg):
    """Fp, Fm = cell averages of polynomials along d: F = Fp_face + Fm_face (the two reconstructions'
    exact values at the face) for UP7 (degree 6), WENO7 (3), WENO5 (2) and NND2 (1)."""
    p = f3d.make_params(9, 8, 7)
    rng = np.random.default_rng(17 + d)
    shp = p.cell_shape()
    idx = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shp), indexing="ij")
    xd = idx[2 - d] - NG
    L = float(shp[2 - d])
    Fp = np.zeros(p.cell_shape(5))
    Fm = np.zeros_like(Fp)
    cp, cm = [], []
    for n in range(5):
        a = np.concatenate([[1.0 + n], 0.3 * (2 * rng.random(deg) - 1) / L ** np.arange(1, deg + 1)])
        b = np.concatenate([[-0.5 - n], 0.3 * (2 * rng.random(deg) - 1) / L ** np.arange(1, deg + 1)])
        Fp[n], Fm[n] = _cell_avg_poly(a, xd), _cell_avg_poly(b, xd)
        cp.append(a)
        cm.append(b)
    phi = np.full(shp, phi0)
    S = np.ones(shp)
    F = np.zeros(p.face_shape(d))
    Oracle(p).advect_comp(d, F, phi, S, Fp, Fm)
    nf = [p.Nxp, p.Nyp, p.Nzp][d] + 1
    xf = np.arange(nf, dtype=np.float64) - 0.5  # face f between cells NG+f-1 and NG+f
    sh = [1, 1, 1]
    sh[2 - d] = nf
    for n in range(5):
        ex = sum(a * xf ** kk for kk, a in enumerate(cp[n])) + sum(b * xf ** kk for kk, b in enumerate(cm[n]))
        ex = np.broadcast_to(ex.reshape(sh), F.shape[1:])
        assert np.abs(F[n] - ex).max() < 1e-10 * np.abs(ex).max(), (n, phi0)
